"""The oracle at run-time precisions (any p, not only the compiled ones) --
the checker of the mpgfx_rt_* path -- against MPFR / MPC through
oracle/build/pin_mpfr: elementwise mul/add/sub, real and complex AXPY and the
complex product agree value for value (zero signs aside: the oracle follows
the fixed types' +0, the run-time path MPFR's signs, tested on the device).
CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from mpcases import edge_arrays
from mpc_cuda_amd.arrays import nlimbs

pytestmark = pytest.mark.skipif(not O.pin_available(), reason="MPFR driver (oracle/build/pin_mpfr) not built")
PRECS = [1, 2, 3, 24, 53, 63, 65, 100, 113, 200, 333, 500, 777, 1000, 1500, 2000]


def operands(p, n=250, seed=1):
    a = O.fill_ref(p, n, 0x51 + p + seed, 25)
    b = O.fill_ref(p, n, 0x73 + p + seed, 25)
    if p < 4:
        return a, b
    ea, eb, _, _ = edge_arrays(p, seed=seed)
    A = (np.ascontiguousarray(np.concatenate([a[0], ea.limbs], 1)), np.ascontiguousarray(np.concatenate([a[1], ea.se])))
    B = (np.ascontiguousarray(np.concatenate([b[0], eb.limbs], 1)), np.ascontiguousarray(np.concatenate([b[1], eb.se])))
    return A, B


def cplx(x, n):
    return (np.ascontiguousarray(np.stack([x[0][:, :n], x[0][:, n:2 * n]])),
            np.ascontiguousarray(np.stack([x[1][:n], x[1][n:2 * n]])))


@pytest.mark.parametrize("p", PRECS)
def test_oracle_runtime_precision_vs_mpfr(p):
    A, B = operands(p)
    n, N = A[1].shape[0], nlimbs(p)
    for op in ("mul", "add", "sub"):
        assert O.mismatches(O.elementwise_real(p, op, A, B), O.pin(op, p, [A, B], ((N, n), (n,)), [n])) == 0, op
    m = 97
    x, y, a = O.fill_ref(p, 2 * m, 0x11 + p, 30), O.fill_ref(p, 2 * m, 0x22 + p, 30), O.fill_ref(p, 2, 0x33 + p, 5)
    ra = (np.ascontiguousarray(a[0][:, :1]), np.ascontiguousarray(a[1][:1]))
    rx = (np.ascontiguousarray(x[0][:, :m]), np.ascontiguousarray(x[1][:m]))
    ry = (np.ascontiguousarray(y[0][:, :m]), np.ascontiguousarray(y[1][:m]))
    assert O.mismatches(O.axpy_real(p, ra, rx, ry), O.pin("axpy", p, [ra, rx, ry], ((N, m), (m,)), [m])) == 0
    ca, cx, cy = cplx(a, 1), cplx(x, m), cplx(y, m)
    assert O.mismatches(O.axpy_complex(p, ca, cx, cy), O.pin("caxpy", p, [ca, cx, cy], ((2, N, m), (2, m)), [m])) == 0
    assert O.mismatches(O.elementwise_complex(p, "mul", cx, cy), O.pin("cmul", p, [cx, cy], ((2, N, m), (2, m)), [m])) == 0
