"""The exact GEMM's asynchronous mode (a caller workspace of
mpgfx_gemm_workspace_size() bytes: the plan decided on the device, no
synchronisation, no allocation): captured in a HIP graph and replayed,
against the synchronous mode (no workspace: the library sizes and allocates
it after reading the exponent scan) and the oracle, on the MFMA path and the
wide-accumulator path; and the refusal of spreads beyond 2^22 accumulator
bits, reported in the workspace's status word (async) or as the return code
(sync)."""
import ctypes

import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd import _lib
from mpc_cuda_amd.arrays import MPArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sync_mode(pb, A, B, m, n, k, C):
    """the C ABI with no workspace: the synchronous mode"""
    import torch
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return _lib.lib().mpgfx_gemm_real(pb, _lib.GEMM_EXACT, m, n, k, *A.ptrs(), *B.ptrs(), *C.ptrs(), None, 0, s)


@pytest.mark.parametrize("pb,m,n,k,lo,hi", [
    (256, 96, 64, 200, -4, 4),          # MFMA path, G = 17
    (256, 40, 70, 130, -80, 80),        # MFMA path, a smaller group size
    (128, 20, 30, 50, -3000, 3000),     # wide-accumulator path
])
def test_exact_async_matches_sync_and_oracle(gpu, pb, m, n, k, lo, hi):
    import torch
    A = M.random_real(pb, m * k, 61 + pb + k, lo, hi)
    B = M.random_real(pb, k * n, 62 + pb + k, lo, hi)
    Ad, Bd = A.cuda(), B.cuda()
    ra = M.gemm(Ad, Bd, m, n, k, "exact")
    assert M.exact_status() == _lib.OK
    rs = M.empty(pb, m * n)
    assert sync_mode(pb, Ad, Bd, m, n, k, rs) == _lib.OK
    torch.cuda.synchronize()
    assert M.equal(ra, rs) == 0
    assert O.mismatches(ra.pair(), O.gemm(pb, m, n, k, A.pair(), B.pair(), "exact"), ignore_zero_sign=False) == 0


def test_exact_async_graph_replays(gpu):
    """Captured once, replayed three times on inputs rewritten between the
    replays (eager copies and a synchronisation in between): every replay
    equals the eager call on the same inputs."""
    import torch
    pb, m, n, k = 256, 64, 96, 160
    sets = [(M.device_random(pb, m * k, 71 + 2 * i, -8, 8), M.device_random(pb, k * n, 72 + 2 * i, -8, 8))
            for i in range(3)]
    A, B = M.empty(pb, m * k), M.empty(pb, k * n)
    C = M.empty(pb, m * n)
    ws = torch.empty(M.gemm_workspace_size(pb, "exact", m, n, k), dtype=torch.uint8, device="cuda")

    def load(i):
        A.limbs.copy_(sets[i][0].limbs); A.se.copy_(sets[i][0].se)
        B.limbs.copy_(sets[i][1].limbs); B.se.copy_(sets[i][1].se)

    load(0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        M.gemm(A, B, m, n, k, "exact", out=C, workspace=ws)       # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        M.gemm(A, B, m, n, k, "exact", out=C, workspace=ws)
    torch.cuda.synchronize()
    for i in (1, 0, 2):
        load(i)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        want = M.gemm(sets[i][0], sets[i][1], m, n, k, "exact")
        torch.cuda.synchronize()
        assert M.equal(C, want) == 0, i
        assert int(ws[:8].view(torch.int64).item()) == _lib.OK


def _planted_undecided(pb, m, n, k, seed):
    """Value lists of A (m x k) and B (k x n): the even rows of A are x + y
    (k indices 0 and 1, the rest of the row zero) with y half an ulp of x (a
    tie, x odd and even), just below it (all-ones significand, whose low
    digits fall below the cut: the truncated sum crosses the midpoint), just
    above it, minus just below it, or -x (exact zero); columns 1, 4 and 6 of
    B hold 1 in their first two entries; everything else random."""
    from mpcases import unpack
    A = M.random_real(pb, m * k, seed, -4, 4)
    B = M.random_real(pb, k * n, seed + 1, -4, 4)
    Av, Bv = unpack(A), unpack(B)
    one = (1, 1, 1 << (pb - 1), False)
    zero = (1, M.ZERO_EXP, 0, True)
    for j in (1, 4, 6):
        Bv[0 * n + j] = one
        Bv[1 * n + j] = one
    rng = np.random.default_rng(seed + 2)
    top = 1 << (pb - 1)
    for r in range(m // 2):
        i = 2 * r
        e = int(rng.integers(-4, 5))
        X = top | (int(rng.integers(1, 1 << 62)) << 2) | (r & 1)       # odd or even
        case = r % 5
        if case == 0:
            y = (1, e - pb, top)                        # half an ulp: a tie
        elif case == 1:
            y = (1, e - pb - 1, (1 << pb) - 1)          # just below half an ulp
        elif case == 2:
            y = (1, e - pb, top | 1)                    # just above
        elif case == 3:
            y = (-1, e - pb - 1, (1 << pb) - 1)         # minus just below
        else:
            y = (-1, e, X)                              # x - x
        Av[i * k + 0] = (1, e, X, False)
        Av[i * k + 1] = (y[0], y[1], y[2], False)
        for c in range(2, k):
            Av[i * k + c] = zero
    return Av, Bv


def test_exact_async_fixup_of_undecided_elements(gpu):
    """The asynchronous mode leaves the digit pairs far below the rounding
    position out of its MFMA stage (the plan's cut) and recomputes exactly
    every element whose rounding the left-out part could change; planted
    ties, near-ties and exact zeros (_planted_undecided) among random
    elements.  The cut is read from the plan block at the start of the
    workspace (int32 at byte 60); results against the synchronous mode (no
    cut) and the oracle."""
    import torch
    from mpcases import pack
    pb, m, n, k = 256, 20, 8, 64
    Av, Bv = _planted_undecided(pb, m, n, k, 91)
    A2, B2 = pack(Av, pb), pack(Bv, pb)
    Ad, Bd = A2.cuda(), B2.cuda()
    ws = torch.empty(M.gemm_workspace_size(pb, "exact", m, n, k), dtype=torch.uint8, device="cuda")
    ra = M.gemm(Ad, Bd, m, n, k, "exact", workspace=ws)
    torch.cuda.synchronize()
    assert int(ws[:8].view(torch.int64).item()) == _lib.OK
    assert int(ws[60:64].view(torch.int32).item()) > 0          # pairs were left out
    rs = M.empty(pb, m * n)
    assert sync_mode(pb, Ad, Bd, m, n, k, rs) == _lib.OK
    torch.cuda.synchronize()
    assert M.equal(ra, rs) == 0
    ref = O.gemm(pb, m, n, k, A2.pair(), B2.pair(), "exact")
    assert O.mismatches(ra.pair(), ref, ignore_zero_sign=False) == 0


def test_exact_async_fixup_in_graph_replays(gpu):
    """The fix-up of undecided elements inside a captured HIP graph: the
    planted rows (ties, near-ties, exact zeros) captured once and replayed
    three times, with the output overwritten between the replays (an eager
    fill and a synchronisation); every replay equals the eager call and the
    oracle."""
    import torch
    from mpcases import pack
    pb, m, n, k = 256, 20, 8, 64
    Av, Bv = _planted_undecided(pb, m, n, k, 131)
    A2, B2 = pack(Av, pb), pack(Bv, pb)
    A, B = A2.cuda(), B2.cuda()
    C = M.empty(pb, m * n)
    ws = torch.empty(M.gemm_workspace_size(pb, "exact", m, n, k), dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        M.gemm(A, B, m, n, k, "exact", out=C, workspace=ws)       # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        M.gemm(A, B, m, n, k, "exact", out=C, workspace=ws)
    torch.cuda.synchronize()
    want = M.gemm(A, B, m, n, k, "exact")
    ref = O.gemm(pb, m, n, k, A2.pair(), B2.pair(), "exact")
    torch.cuda.synchronize()
    for r in range(3):
        C.limbs.fill_(0x5A5A + r)
        C.se.fill_(12345 + r)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(ws[:8].view(torch.int64).item()) == _lib.OK
        assert int(ws[60:64].view(torch.int32).item()) > 0
        assert M.equal(C, want) == 0, r
        assert O.mismatches(C.pair(), ref, ignore_zero_sign=False) == 0, r


def test_complex_exact_fixup_of_undecided_elements(gpu):
    """The complex exact GEMM (two real exact GEMMs over 2k terms in the
    asynchronous mode) on the planted rows as real parts, with zero
    imaginary parts on those rows: the ties and near-ties land on Re; against
    the oracle."""
    from mpcases import pack, unpack
    pb, m, n, k = 256, 20, 8, 64
    Av, Bv = _planted_undecided(pb, m, n, k, 95)
    Ai = M.random_real(pb, m * k, 97, -4, 4)
    Bi = M.random_real(pb, k * n, 98, -4, 4)
    Aiv, Biv = unpack(Ai), unpack(Bi)
    zero = (1, M.ZERO_EXP, 0, True)
    for i in range(0, m, 2):
        for c in range(k):
            Aiv[i * k + c] = zero
    Ar, Air = pack(Av, pb), pack(Aiv, pb)
    Br, Bir = pack(Bv, pb), pack(Biv, pb)
    A = MPArray(pb, np.ascontiguousarray(np.stack([Ar.limbs, Air.limbs])),
                np.ascontiguousarray(np.stack([Ar.se, Air.se])), True)
    B = MPArray(pb, np.ascontiguousarray(np.stack([Br.limbs, Bir.limbs])),
                np.ascontiguousarray(np.stack([Br.se, Bir.se])), True)
    r = M.gemm(A.cuda(), B.cuda(), m, n, k, "exact")
    ref = O.gemm(pb, m, n, k, A.pair(), B.pair(), "cexact")
    assert O.mismatches(r.pair(), ref, ignore_zero_sign=False) == 0


def test_exact_span_beyond_cap_refused(gpu):
    """Exponent spreads of about 2^21 in both operands need more than 2^22
    accumulator bits: the synchronous mode returns EUNSUPPORTED, the
    asynchronous one sets the status word and writes +0."""
    import torch
    pb, m, n, k = 64, 4, 3, 8
    A = M.random_real(pb, m * k, 81, -(1 << 21), 1 << 21).cuda()
    B = M.random_real(pb, k * n, 82, -(1 << 21), 1 << 21).cuda()
    C = M.empty(pb, m * n)
    assert sync_mode(pb, A, B, m, n, k, C) == _lib.EUNSUPPORTED
    C.se.fill_(12345)
    M.gemm(A, B, m, n, k, "exact", out=C)
    assert M.exact_status() == _lib.EUNSUPPORTED
    h = C.numpy()
    assert np.all((h.se >> 1) == M.ZERO_EXP) and np.all(h.limbs == 0)


def test_complex_exact_span_beyond_cap_refused(gpu):
    """The complex exact GEMM runs its two real GEMMs in the asynchronous mode
    and reads their status words at its end: a spread beyond 2^22
    accumulator bits comes back as EUNSUPPORTED from the call."""
    pb, m, n, k = 64, 3, 4, 6
    A = M.random_complex(pb, m * k, 83, -(1 << 21), 1 << 21).cuda()
    B = M.random_complex(pb, k * n, 84, -(1 << 21), 1 << 21).cuda()
    with pytest.raises(M.MPGFXError):
        M.gemm(A, B, m, n, k, "exact")
