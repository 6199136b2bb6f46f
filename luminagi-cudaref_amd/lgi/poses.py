"""Camera pose samplers.

* orbit_poses: poses on a horizontal orbit around the scene with look-at jitter (the
  multi-pose workload: 15 m orbit, +-2 m jitter, spp budgets log-uniform in [8, 256]).
* random_camera_around: the reference dataset sampler (ml/gen_dataset.py:16-30): position in a
  disc of radius 2.5 m around an anchor (+-0.4 m in z), yaw +-30 deg, pitch +-15 deg, target one
  metre ahead. Uses Python's `random` module like the reference, so a seeded run reproduces
  the reference's pose sequence for the same anchor.
"""
import math
import random

import numpy as np

from .tracer import camera_look_at


def orbit_poses(n, seed=0, radius=15.0, center=(0.0, 0.0, 6.0), height=(3.0, 12.0), jitter=2.0,
                spp_range=(8, 256)):
    """Returns [(pos, target, spp)] with positions on the orbit and log-uniform spp budgets."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        ang = 2.0 * math.pi * (i + rng.uniform(0.0, 1.0)) / n
        pos = (center[0] + radius * math.cos(ang), center[1] + radius * math.sin(ang), rng.uniform(*height))
        tgt = tuple(float(c + rng.uniform(-jitter, jitter)) for c in center)
        lo, hi = math.log(spp_range[0]), math.log(spp_range[1])
        spp = int(round(math.exp(rng.uniform(lo, hi))))
        out.append((tuple(float(p) for p in pos), tgt, spp))
    return out


def random_camera_around(anchor, radius_xy=2.5, radius_z=0.4, yaw_deg=30, pitch_deg=15):
    """ml/gen_dataset.py:16-30 (uses the global `random` state)."""
    ax, ay, az = anchor
    r = random.uniform(0.0, radius_xy)
    ang = random.uniform(0.0, 2.0 * math.pi)
    pos = (ax + r * math.cos(ang), ay + r * math.sin(ang), az + random.uniform(-radius_z, radius_z))
    yaw = math.radians(random.uniform(-yaw_deg, yaw_deg))
    pitch = math.radians(random.uniform(-pitch_deg, pitch_deg))
    fx = math.cos(pitch) * math.cos(yaw)
    fy = math.cos(pitch) * math.sin(yaw)
    fz = math.sin(pitch)
    return pos, (pos[0] + fx, pos[1] + fy, pos[2] + fz)


def camera_for(base_cam, pos, target):
    """Copy of base_cam looking from pos at target (overwrite_camera, src/main.cpp:29-59)."""
    from ._abi import Camera
    c = Camera.from_buffer_copy(base_cam)
    camera_look_at(c, pos, target)
    return c
