"""Loading of the committed golden fixtures (tests/golden/pb*.npz)."""
import os

import numpy as np

from mpc_cuda_amd.arrays import MPArray

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PBS = [32, 64, 96, 128, 160, 192, 256, 288, 320, 384, 512, 1024, 1056, 2048]


def load(pb):
    return dict(np.load(os.path.join(GOLDEN, f"pb{pb}.npz")))


def arr(g, key, pb, cplx=False):
    return MPArray(pb, np.ascontiguousarray(g[key + "_limbs"]), np.ascontiguousarray(g[key + "_se"]), cplx)


def pair(g, key):
    return np.ascontiguousarray(g[key + "_limbs"]), np.ascontiguousarray(g[key + "_se"])
