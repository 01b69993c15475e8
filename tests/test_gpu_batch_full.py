"""The batched mixed workload at its full size: 10,000 AXPY / dot / GEMV
instances (40 / 40 / 20 %, half complex, lengths log-uniform in [16, 65536],
GEMV rows in [16, 2048], precision in {64, 128, 256, 512, 1024}, exponents in
[-8, 8]; mpc_cuda_amd.workloads.batch_mix) in one graph launch, with the fixed
mapping and with the tuned one.  A documented sample is checked bit for bit
(limbs, sign and exponent of every output value) against the oracle:
  - the 20 instances with the longest fold chains,
  - one instance of every (kind, precision, real/complex) class,
  - random instances up to 320 in all.
Outputs are clobbered before the launch, so every checked value was written
by it."""
import concurrent.futures as cf
import os
import random
import time

import pytest

from mpc_cuda_amd import batch as B
from mpc_cuda_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SAMPLE = 320


@pytest.fixture(scope="module")
def mix(gpu):
    import torch
    insts, tasks = W.batch_mix(10000, seed=5)
    torch.cuda.synchronize()
    return insts, tasks


def sample(insts, k=SAMPLE, seed=11):
    idx = set(sorted(range(len(insts)), key=lambda i: -insts[i].chain)[:20])
    seen = {}
    for i, x in enumerate(insts):
        seen.setdefault((x.kind, x.pb, x.cplx), i)
    idx |= set(seen.values())
    rng = random.Random(seed)
    rest = [i for i in range(len(insts)) if i not in idx]
    idx |= set(rng.sample(rest, k - len(idx)))
    return sorted(idx), len(seen)


def oracle_of(x):
    if x.kind == B.AXPY:
        a, xx, y, _ = x.arrays
        f = O.axpy_complex if x.cplx else O.axpy_real
        return f(x.pb, a.pair(), xx.pair(), y.pair())
    A, v = x.arrays[0], x.arrays[1]
    return O.gemm(x.pb, x.R, 1, x.L, A.pair(), v.pair(), "cperop" if x.cplx else "perop")


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("tuning", [False, True], ids=["fixed", "tuned"])
def test_batch_10000_sampled_vs_oracle(mix, tuning, tmp_path):
    import torch
    insts, tasks = mix
    idx, nclasses = sample(insts)
    assert nclasses == 30 and len(idx) == SAMPLE
    for i in idx:
        insts[i].out.limbs.fill_(-1)
        insts[i].out.se.fill_(12345)
    t0 = time.time()
    b = B.Batch(tasks, tuning=tuning, cache=str(tmp_path / "tune.json") if tuning else None)
    b.run()
    torch.cuda.synchronize()
    print(f"batch of {len(insts)} ({'tuned' if tuning else 'fixed'}): created and run in {time.time() - t0:.1f} s",
          flush=True)
    got = {i: insts[i].out.pair() for i in idx}
    b.close()
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        refs = dict(zip(idx, ex.map(lambda i: oracle_of(insts[i]), idx)))
    bad = [(i, insts[i].kind, insts[i].pb, insts[i].cplx, insts[i].R, insts[i].L)
           for i in idx if O.mismatches(got[i], refs[i], ignore_zero_sign=False)]
    print(f"oracle over {len(idx)} sampled instances in {time.time() - t0:.1f} s; "
          f"longest chain {max(insts[i].chain for i in idx)}", flush=True)
    assert not bad, bad[:10]
