"""The DPP cooperative fold (mpc_cuda_amd/csrc/coop_dpp.h) run in a host
emulation of a 64-lane wavefront (tests/coop_emul: one thread per lane, the
cross-lane operations with gfx950's lane mappings) on fold groups laid out as
the GEMV and the batch lay them out -- R = 64/G chains per wavefront, chains
shorter than the wavefront's longest (their extra steps run and are
discarded), zero-length chains, a wavefront whose chains are all empty, and a
partial last wavefront -- against the oracle's per-op fold
(acc = RNDN(acc + p) with cu_fadd's zero rules), bit for bit.  The step
operands cover the structured edge pairs (ties, carry-out, exact and massive
cancellation, far operands with sticky borrow, zeros) and random values with
equal, near and far exponents."""
import ctypes
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from mpcases import edge_pairs, pack
import mpc_cuda_amd as M

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "coop_emul", "libcoop_emul.so")
u64p, i64p = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64)


@pytest.fixture(scope="module")
def emul():
    if not os.path.exists(LIB):
        pytest.fail("tests/coop_emul/libcoop_emul.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.emul_fold.argtypes = [ctypes.c_int, ctypes.c_int64, i64p, i64p, ctypes.c_int64, i64p, u64p, u64p, i64p]
    return L


def chains_for(pb, seed, nchains):
    """list of chains, each a list of value tuples"""
    rng = random.Random(seed)
    A, B = edge_pairs(pb, seed=seed, n_random=24)
    chains = [[a, b] for a, b in zip(A, B)]                 # every edge pair as a 2-step fold
    vals = M.random_real(pb, 4000, seed, -60, 60)
    from mpcases import unpack
    pool = unpack(vals)
    N = (pb + 63) // 64
    while len(chains) < nchains:
        L = rng.choice([0, 0, 1, 3, 17, 40, 64])
        ch = []
        for s in range(L):
            v = rng.choice(pool)
            r = rng.random()
            if ch and r < 0.15:
                p = ch[-1]
                v = (-p[0], p[1], p[2], p[3])                 # exact cancellation of the last term
            elif ch and r < 0.3:
                p = rng.choice(ch)
                v = (-p[0], p[1], p[2] ^ (1 << (64 * N - pb)), p[3]) if not p[3] else p   # near cancellation
            elif r < 0.4:
                v = (v[0], rng.randint(-3, 3), v[2], v[3])   # equal / near exponents
            elif r < 0.45:
                v = (1, O.ZERO_EXP, 0, True)
            ch.append(v)
        chains.append(ch)
    rng.shuffle(chains)
    return chains


def layout(pb, chains, G):
    R = 64 // G
    N = (pb + 63) // 64
    n = len(chains)
    nw = (n + R - 1) // R
    lens = np.array([len(c) for c in chains], np.int64)
    wave_max = np.array([max([len(chains[c]) for c in range(q * R, min(n, q * R + R))] + [0]) for q in range(nw)],
                        np.int64)
    if nw > 1:                                             # one wavefront with only empty chains
        for c in range(R, min(n, 2 * R)):
            chains[c] = []
        lens[R:min(n, 2 * R)] = 0
        wave_max[1] = 0
    base = np.concatenate([[0], np.cumsum(wave_max * R)])[:-1]
    rec0 = np.array([base[c // R] + c % R for c in range(n)], np.int64)
    total = int(max(1, (wave_max * R).sum()))
    recs = np.zeros((total, N + 1), np.uint64)
    for c, ch in enumerate(chains):
        for s, (sg, e, Mv, z) in enumerate(ch):
            r = rec0[c] + s * R
            for l in range(N):
                recs[r, l] = (Mv >> (64 * l)) & 0xFFFFFFFFFFFFFFFF
            recs[r, N] = np.uint64(((O.ZERO_EXP if z else e) * 2 + (1 if sg < 0 else 0)) & 0xFFFFFFFFFFFFFFFF)
    return lens, rec0, R, wave_max, recs


def oracle_fold(pb, chains):
    n = len(chains)
    L = max([len(c) for c in chains] + [0])
    acc = pack([(1, O.ZERO_EXP, 0, True)] * n, pb)
    al, ase = np.ascontiguousarray(acc.limbs), np.ascontiguousarray(acc.se)
    for s in range(L):
        live = np.array([s < len(c) for c in chains])
        p = pack([c[s] if s < len(c) else (1, O.ZERO_EXP, 0, True) for c in chains], pb)
        nl, ns = O.elementwise_real(pb, "add", (al, ase), (np.ascontiguousarray(p.limbs), np.ascontiguousarray(p.se)))
        al = np.where(live[None, :], nl, al)
        ase = np.where(live, ns, ase)
    return np.ascontiguousarray(al), np.ascontiguousarray(ase)


@pytest.mark.parametrize("pb", [32, 64, 96, 128, 192, 256, 320, 512, 1024, 1056, 2048])
def test_dpp_fold_emulated_vs_oracle(emul, pb):
    G = emul.emul_lanes(pb)
    R = 64 // G
    chains = chains_for(pb, 100 + pb, 5 * R + R // 2)       # 5.5 wavefronts
    lens, rec0, step, wave_max, recs = layout(pb, chains, G)
    n = len(chains)
    ol, ose = O.empty_real(pb, n)
    rc = emul.emul_fold(pb, n, lens.ctypes.data_as(i64p), rec0.ctypes.data_as(i64p), step,
                        wave_max.ctypes.data_as(i64p), recs.ctypes.data_as(u64p), ol.ctypes.data_as(u64p),
                        ose.ctypes.data_as(i64p))
    assert rc == 0
    ref = oracle_fold(pb, chains)
    assert O.mismatches((ol, ose), ref, ignore_zero_sign=False) == 0
