"""Run-time precision path on the device (mpgfx_rt_elementwise_real,
mpgfx_rt_axpy_real, mpgfx_rt_axpy_complex) against the oracle at precisions
that are not compiled in (and some that are), value for value; MPFR's zero
signs on constructed zeros; and agreement with the fixed-precision kernels at
a compiled precision.  The oracle at these precisions is pinned to MPFR/MPC
in tests/test_rt_oracle.py."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray, nlimbs
from oracle import oracle as O
from mpcases import pack, unpack
from test_rt_oracle import operands, cplx

pytestmark = pytest.mark.gpu
PRECS = [1, 2, 24, 53, 64, 65, 100, 113, 128, 200, 256, 333, 500, 512, 777, 1000, 1024, 1500, 2000, 2048]


def dev(pair, p, c=False):
    return MPArray(p, np.ascontiguousarray(pair[0]), np.ascontiguousarray(pair[1]), c).cuda()


def by_value(gpu_arr, oracle_pair):
    return O.mismatches(gpu_arr.pair(), oracle_pair, ignore_zero_sign=True)


@pytest.mark.parametrize("p", PRECS)
def test_rt_elementwise_and_axpy_vs_oracle(gpu, p):
    A, B = operands(p, n=300, seed=3)
    dA, dB = dev(A, p), dev(B, p)
    for op in ("mul", "add", "sub"):
        assert by_value(M.rt_elementwise(op, dA, dB), O.elementwise_real(p, op, A, B)) == 0, op
    m = 1031                                   # ragged length
    x, y, a = O.fill_ref(p, 2 * m, 0x41 + p, 30), O.fill_ref(p, 2 * m, 0x42 + p, 30), O.fill_ref(p, 2, 0x43 + p, 5)
    ra = (np.ascontiguousarray(a[0][:, :1]), np.ascontiguousarray(a[1][:1]))
    rx = (np.ascontiguousarray(x[0][:, :m]), np.ascontiguousarray(x[1][:m]))
    ry = (np.ascontiguousarray(y[0][:, :m]), np.ascontiguousarray(y[1][:m]))
    assert by_value(M.rt_axpy(dev(ra, p), dev(rx, p), dev(ry, p)), O.axpy_real(p, ra, rx, ry)) == 0
    ca, cx, cy = cplx(a, 1), cplx(x, m), cplx(y, m)
    assert by_value(M.rt_axpy(dev(ca, p, True), dev(cx, p, True), dev(cy, p, True)),
                    O.axpy_complex(p, ca, cx, cy)) == 0


@pytest.mark.parametrize("p", [53, 300, 1024])
def test_rt_zero_signs_follow_mpfr(gpu, p):
    top = 1 << (p - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(p) - p), False)
    pz, nz = (1, O.ZERO_EXP, 0, True), (-1, O.ZERO_EXP, 0, True)
    x = v(top | 1, 3)
    A = [v(top, 2, -1), pz, nz, nz, pz, x, nz, pz]
    B = [pz, v(top, 2, -1), v(top, 2, -1), nz, nz, x, pz, pz]
    dA, dB = pack(A, p).cuda(), pack(B, p).cuda()
    sign_of = lambda arr: [(-1 if (int(s) & 1) else 1) for s in arr.numpy().se]
    zero_of = lambda arr: [int(s) >> 1 == O.ZERO_EXP for s in arr.numpy().se]
    mul = M.rt_elementwise("mul", dA, dB)
    assert all(zero_of(mul)[k] for k in (0, 1, 2, 3, 4, 6, 7))
    assert [sign_of(mul)[k] for k in (0, 1, 2, 3, 4, 6, 7)] == [-1, -1, 1, 1, -1, -1, 1]
    add = M.rt_elementwise("add", dA, dB)
    assert [zero_of(add)[k] for k in (3, 4, 6, 7)] == [True] * 4
    assert [sign_of(add)[k] for k in (3, 4, 6, 7)] == [-1, 1, 1, 1]       # -0 + -0 = -0, others +0
    sub = M.rt_elementwise("sub", dA, dB)
    assert zero_of(sub)[5] and sign_of(sub)[5] == 1                       # x - x = +0
    assert [sign_of(sub)[k] for k in (3, 4, 6, 7)] == [1, 1, -1, 1]       # -0 - +0 = -0


@pytest.mark.parametrize("p", [256, 1024])
def test_rt_matches_fixed_precision_kernels(gpu, p):
    a = M.random_real(p, 1, 5, -8, 8).cuda()
    x = M.random_real(p, 5000, 6, -20, 20).cuda()
    y = M.random_real(p, 5000, 7, -20, 20).cuda()
    r1, r2 = M.rt_axpy(a, x, y).numpy(), M.axpy(a, x, y).numpy()
    assert np.array_equal(r1.limbs, r2.limbs) and np.array_equal(r1.se, r2.se)
    ca = M.random_complex(p, 1, 8, -8, 8).cuda()
    cx = M.random_complex(p, 3000, 9, -20, 20).cuda()
    cy = M.random_complex(p, 3000, 10, -20, 20).cuda()
    r1, r2 = M.rt_axpy(ca, cx, cy).numpy(), M.axpy(ca, cx, cy).numpy()
    assert np.array_equal(r1.limbs, r2.limbs) and np.array_equal(r1.se, r2.se)
