"""Batched execution (one hipGraph launch for many AXPY / dot / GEMV instances
of mixed precision, real and complex) against the oracle instance by instance,
over several graph replays, with the fixed mapping and the tuned one."""
import os
import random

import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd import batch as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_tasks(seed, n_tasks, max_len, pbs=(64, 128, 256, 512, 1024)):
    rng = random.Random(seed)
    specs = []
    for t in range(n_tasks):
        kind = rng.choices([B.AXPY, B.DOT, B.GEMV], [0.4, 0.4, 0.2])[0]
        pb = rng.choice(pbs)
        cplx = rng.random() < 0.5
        n = int(round(2 ** rng.uniform(np.log2(1), np.log2(max_len))))
        if rng.random() < 0.05:
            n = 0                                   # empty instances
        m = int(round(2 ** rng.uniform(0, 5))) if kind == B.GEMV else 1
        specs.append((kind, pb, cplx, m, n, 1000 + t))
    return specs


def build(specs):
    tasks, refs = [], []
    for kind, pb, cplx, m, n, sd in specs:
        gen = lambda cnt, s: M.random_array(pb, cnt, s, -8, 8, cplx)
        if kind == B.AXPY:
            a, x, y = gen(1, sd), gen(n, sd + 1), gen(n, sd + 2)
            out = M.empty(pb, n, cplx)
            tasks.append(B.axpy(a.cuda(), x.cuda(), y.cuda(), out))
            ref = (O.axpy_complex if cplx else O.axpy_real)(pb, a.pair(), x.pair(), y.pair())
        elif kind == B.DOT:
            x, y = gen(n, sd), gen(n, sd + 1)
            out = M.empty(pb, 1, cplx)
            tasks.append(B.dot(x.cuda(), y.cuda(), out))
            ref = O.gemm(pb, 1, 1, n, x.pair(), y.pair(), "cperop" if cplx else "perop")
        else:
            A, x = gen(m * n, sd), gen(n, sd + 1)
            out = M.empty(pb, m, cplx)
            tasks.append(B.gemv(A.cuda(), x.cuda(), m, n, out))
            ref = O.gemm(pb, m, 1, n, A.pair(), x.pair(), "cperop" if cplx else "perop")
        refs.append((out, ref))
    return tasks, refs


def check(refs):
    bad = 0
    for out, ref in refs:
        bad += O.mismatches(out.pair(), ref, ignore_zero_sign=False)
    return bad


@pytest.mark.parametrize("tuning", [0, 1])
def test_batch_mixed_vs_oracle(gpu, tuning, tmp_path):
    import torch
    specs = make_tasks(7 + tuning, 160, 3000)
    tasks, refs = build(specs)
    cache = str(tmp_path / "tune.json") if tuning else None
    b = B.Batch(tasks, tuning=bool(tuning), cache=cache)
    info = b.info()
    assert info["classes"] >= 8
    for rep in range(3):
        for out, _ in refs:                       # clobber outputs between replays
            out.limbs.fill_(-1)
            out.se.fill_(12345)
        b.run()
        torch.cuda.synchronize()
        assert check(refs) == 0, rep
        _ = M.axpy(*(M.random_real(256, 1, 1).cuda(), M.random_real(256, 64, 2).cuda(), M.random_real(256, 64, 3).cuda()))
    if tuning:
        assert os.path.exists(cache)
        b2 = B.Batch(tasks, tuning=True, cache=cache)        # cached mapping
        b2.run()
        torch.cuda.synchronize()
        assert check(refs) == 0
        b2.close()
    b.close()


def test_batch_long_chains(gpu):
    """A few long dots / GEMV rows (the fold critical path) next to short work."""
    import torch
    specs = [(B.DOT, 256, False, 1, 20000, 1), (B.DOT, 1024, True, 1, 3000, 2),
             (B.GEMV, 512, False, 70, 2000, 3), (B.AXPY, 128, True, 1, 50000, 4),
             (B.GEMV, 64, True, 3, 9000, 5), (B.DOT, 64, False, 1, 1, 6)]
    tasks, refs = build(specs)
    b = B.Batch(tasks)
    b.run()
    torch.cuda.synchronize()
    assert check(refs) == 0


@pytest.mark.parametrize("pb,cplx", [(64, False), (96, True), (256, False), (1024, True), (2048, False)])
def test_device_generator_matches_host(gpu, pb, cplx):
    n = 10007
    d = M.device_random(pb, n, 99, -30, 30, cplx)
    h = M.random_array(pb, n, 99, -30, 30, cplx)
    assert M.equal(d, h) == 0


@pytest.mark.parametrize("tuning", [0, 1])
def test_batch_inplace_axpy(gpu, tuning, tmp_path):
    """y = a*x + y in place (out is y) next to a dot and a GEMV: the tuner runs
    the batch's products many times while it measures, yet the first real run
    must see the original y; a second run applies the AXPY again."""
    import torch
    pb, n = 256, 5000
    a, x = M.random_real(pb, 1, 1), M.random_real(pb, n, 2)
    y0 = M.random_real(pb, n, 3)
    y = y0.cuda()
    u, v = M.random_complex(512, 3000, 4), M.random_complex(512, 3000, 5)
    dout = M.empty(512, 1, True)
    A, xv = M.random_real(128, 40 * 700, 6), M.random_real(128, 700, 7)
    gout = M.empty(128, 40)
    tasks = [B.axpy(a.cuda(), x.cuda(), y, y), B.dot(u.cuda(), v.cuda(), dout),
             B.gemv(A.cuda(), xv.cuda(), 40, 700, gout)]
    b = B.Batch(tasks, tuning=bool(tuning), cache=str(tmp_path / "t.json") if tuning else None)
    b.run()
    torch.cuda.synchronize()
    r1 = O.axpy_real(pb, a.pair(), x.pair(), y0.pair())
    assert O.mismatches(y.pair(), r1, ignore_zero_sign=False) == 0
    assert O.mismatches(dout.pair(), O.gemm(512, 1, 1, 3000, u.pair(), v.pair(), "cperop"), ignore_zero_sign=False) == 0
    assert O.mismatches(gout.pair(), O.gemm(128, 40, 1, 700, A.pair(), xv.pair(), "perop"), ignore_zero_sign=False) == 0
    b.run()
    torch.cuda.synchronize()
    r2 = O.axpy_real(pb, a.pair(), x.pair(), r1)
    assert O.mismatches(y.pair(), r2, ignore_zero_sign=False) == 0
    b.close()


def test_batch_task_shape_checks(gpu):
    """task wrappers reject arrays of the wrong length, precision or kind"""
    x, y = M.random_real(256, 10, 1).cuda(), M.random_real(256, 10, 2).cuda()
    with pytest.raises(ValueError):
        B.dot(x, M.random_real(256, 9, 3).cuda(), M.empty(256, 1))
    with pytest.raises(ValueError):
        B.axpy(M.random_real(256, 1, 4).cuda(), x, y, M.empty(256, 11))
    with pytest.raises(ValueError):
        B.gemv(M.random_real(256, 20, 5).cuda(), x, 3, 10, M.empty(256, 3))
    with pytest.raises(ValueError):
        B.dot(x, M.random_real(128, 10, 6).cuda(), M.empty(256, 1))
