"""train.py — render (8-spp noisy, 256-spp clean) pairs with the project's tracer, train the
denoiser on its own HIP kernels, save the best checkpoint and evaluate time-to-40-dB.
Entry point of the reference's ml/train.py; see lgi.train for the recipe.

    python luminagi-cudaref_amd/ml/train.py [--poses 50] [--steps 2000] [--time-budget S]
                                             [--out denoiser.pt] [--report eval.json]
"""
import pathlib
import sys

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent.parent))

from lgi.train import main  # noqa: E402

if __name__ == "__main__":
    main()
