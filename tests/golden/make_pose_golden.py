"""Writes pose_golden.json in this folder: the 50 camera poses (position, target) of the
reference's committed dataset batch file, ml/data/batch.txt, which ml/gen_dataset.py wrote with
random.seed(42) (ml/gen_dataset.py:44, 46-65). Each pose appears twice there (noisy and clean
job); the file keeps one copy per pose, in order, as the 4-decimal numbers the batch file holds.

Run in the build container, where /root/reference is mounted:

    python tests/golden/make_pose_golden.py
"""
import json
import pathlib

SRC = pathlib.Path("/root/reference/ml/data/batch.txt")
OUT = pathlib.Path(__file__).resolve().parent / "pose_golden.json"


def main():
    poses = []
    for line in SRC.read_text().splitlines():
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        pos, tgt = [float(x) for x in f[3:6]], [float(x) for x in f[6:9]]
        if "/noisy/" in f[0]:
            poses.append({"pos": pos, "target": tgt})
    rows = ",\n".join("  " + json.dumps(p) for p in poses)
    OUT.write_text('{"seed": 42, "source": "ml/data/batch.txt", "poses": [\n' + rows + "\n]}\n")
    print(f"wrote {OUT} ({len(poses)} poses)")


if __name__ == "__main__":
    main()
