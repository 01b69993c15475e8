"""The batched mixed workload of the benchmark: many AXPY / dot / GEMV
instances of uneven length, mixed precision, real and complex, on inputs
generated on the GPU.

Default mix (10,000 instances): kind AXPY / dot / GEMV at 40 / 40 / 20 %, half
complex, lengths log-uniform in [16, 65536], GEMV rows log-uniform in
[16, 2048] with the matrix capped at 2^18 values (columns reduced), precision
uniform in {64, 128, 256, 512, 1024}, exponents uniform in [-8, 8].  Dot and
GEMV are the reference's per-operation k-order fold
(tools/linalg_fast.cu:37-51).
"""
import math
import random

from . import batch as B
from .arrays import empty
from .gen import device_random


class Instance:
    """One instance: its kind, precision, complex flag, rows R and length L
    (AXPY and dot: R = 1), and its device arrays (inputs, then the output)."""
    __slots__ = ("kind", "pb", "cplx", "R", "L", "arrays", "seed")

    def __init__(self, kind, pb, cplx, R, L, arrays, seed):
        self.kind, self.pb, self.cplx, self.R, self.L, self.arrays, self.seed = kind, pb, cplx, R, L, arrays, seed

    @property
    def out(self):
        return self.arrays[-1]

    @property
    def chain(self):
        """length of the dependent fold chain (0 for AXPY)"""
        return 0 if self.kind == B.AXPY else self.L

    def elements(self):
        return self.R * self.L

    def word_products(self):
        return self.elements() * (self.pb // 32) ** 2 * (4 if self.cplx else 1)


def batch_mix(n_tasks=10000, seed=5, pbs=(64, 128, 256, 512, 1024), max_len=65536, max_matrix=1 << 18,
              emin=-8, emax=8):
    """Instances and their batch task descriptors (mpc_cuda_amd.batch)."""
    rng = random.Random(seed)

    def logu(a, b):
        return int(round(2 ** rng.uniform(math.log2(a), math.log2(b))))

    insts, tasks = [], []
    for t in range(n_tasks):
        kind = rng.choices([B.AXPY, B.DOT, B.GEMV], [0.4, 0.4, 0.2])[0]
        pb = rng.choice(pbs)
        cplx = rng.random() < 0.5
        L = logu(16, max_len)
        R = 1
        if kind == B.GEMV:
            R = logu(16, 2048)
            L = max(1, min(L, max_matrix // R))
        sd = 7919 * t + 13
        g = lambda cnt, s: device_random(pb, cnt, s, emin, emax, cplx)
        if kind == B.AXPY:
            arrs = (g(1, sd), g(L, sd + 1), g(L, sd + 2), empty(pb, L, cplx))
            tasks.append(B.axpy(*arrs))
        elif kind == B.DOT:
            arrs = (g(L, sd), g(L, sd + 1), empty(pb, 1, cplx))
            tasks.append(B.dot(*arrs))
        else:
            arrs = (g(R * L, sd), g(L, sd + 1), empty(pb, R, cplx))
            tasks.append(B.gemv(arrs[0], arrs[1], R, L, arrs[2]))
        insts.append(Instance(kind, pb, cplx, R, L, arrs, sd))
    return insts, tasks
