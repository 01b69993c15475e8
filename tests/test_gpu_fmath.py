"""Division and square root on the device (mpgfx_elementwise_real with
MPGFX_DIV, mpgfx_sqrt_real; cu_fdiv / cu_sqrt of include/mpc_cuda/cu_fmath.cuh)
against the oracle, bit for bit: the reference generator's values, the
structured edge cases, zero and negative operands, and representable results.
The oracle's restatement is pinned to the reference build and to MPFR in
tests/test_fmath.py."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray, nlimbs
from oracle import oracle as O
import goldens
from mpcases import pack
from test_fmath import operands

pytestmark = pytest.mark.gpu


def dev(pair, pb):
    return MPArray(pb, np.ascontiguousarray(pair[0]), np.ascontiguousarray(pair[1])).cuda()


def same(gpu_arr, oracle_pair):
    return O.mismatches(gpu_arr.pair(), oracle_pair, ignore_zero_sign=False)


@pytest.mark.parametrize("pb", goldens.PBS)
def test_div_sqrt_vs_oracle(gpu, pb):
    A, B = operands(pb, n=700, seed=21)
    assert same(M.elementwise("div", dev(A, pb), dev(B, pb)), O.elementwise_real(pb, "div", A, B)) == 0
    assert same(M.sqrt(dev(A, pb)), O.sqrt_real(pb, A)) == 0


@pytest.mark.parametrize("pb", [64, 1024])
def test_div_sqrt_special_operands(gpu, pb):
    """+0 for a zero dividend, a zero divisor, a zero or negative radicand;
    exact results where representable."""
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    z = (1, O.ZERO_EXP, 0, True)
    num = pack([z, v(top, 1), v(3 << (pb - 2), 3), v(top | 1, 5, -1)], pb)
    den = pack([v(top, 1), z, v(3 << (pb - 2), 2), v(top | 1, 5, -1)], pb)
    q = M.elementwise("div", num.cuda(), den.cuda())
    assert same(q, O.elementwise_real(pb, "div", num.pair(), den.pair())) == 0
    rad = pack([z, v(top, 3, -1), v(top, 3), v(9 << (pb - 4), 0), v(top | 1, -7)], pb)
    r = M.sqrt(rad.cuda())
    assert same(r, O.sqrt_real(pb, rad.pair())) == 0
    h = r.numpy()
    assert (h.se[:2] >> 1 == O.ZERO_EXP).all() and (h.se[:2] & 1 == 0).all()


def test_div_rejects_complex_and_sqrt_rejects_complex(gpu):
    a = M.random_complex(256, 8, 1).cuda()
    with pytest.raises(ValueError):
        M.elementwise("div", a, a)
    with pytest.raises(ValueError):
        M.sqrt(a)
