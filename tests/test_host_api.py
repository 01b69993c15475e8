"""The drop-in template API (include/mpc_cuda/*.cuh) compiled for the host by
g++ (tests/host/libcufp_host.so) agrees bit for bit with the oracle -- the same
engine the HIP kernels instantiate, checked on the CPU."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
import goldens
from mpcases import edge_arrays

HOSTLIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libcufp_host.so")
u64p, i64p = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64)


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOSTLIB):
        pytest.fail("tests/host/libcufp_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(HOSTLIB)
    for f in (L.host_elementwise_real, L.host_elementwise_complex):
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + [u64p, i64p] * 3
    L.host_roundtrip_double.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    L.host_roundtrip_double.restype = ctypes.c_double
    return L


def P(a, t):
    return a.ctypes.data_as(t)


@pytest.mark.parametrize("pb", goldens.PBS)
def test_host_template_api_vs_oracle(host, pb):
    M = 4000
    l, s = O.fill_ref(pb, 2 * M, 0xABC + pb, -1)
    ea, eb, _, _ = edge_arrays(pb, seed=9)
    a = (np.ascontiguousarray(np.concatenate([l[:, 0::2], ea.limbs], 1)), np.ascontiguousarray(np.concatenate([s[0::2], ea.se])))
    b = (np.ascontiguousarray(np.concatenate([l[:, 1::2], eb.limbs], 1)), np.ascontiguousarray(np.concatenate([s[1::2], eb.se])))
    n = a[1].shape[0]
    for k, op in enumerate(("mul", "add", "sub")):
        hl, hs = O.empty_real(pb, n)
        assert host.host_elementwise_real(pb, k, n, P(a[0], u64p), P(a[1], i64p), P(b[0], u64p), P(b[1], i64p),
                                          P(hl, u64p), P(hs, i64p)) == 0
        assert O.mismatches(O.elementwise_real(pb, op, a, b), (hl, hs), ignore_zero_sign=False) == 0, op
    h = n // 2
    ca = (np.ascontiguousarray(np.stack([a[0][:, :h], b[0][:, h:2 * h]])), np.ascontiguousarray(np.stack([a[1][:h], b[1][h:2 * h]])))
    cb = (np.ascontiguousarray(np.stack([b[0][:, :h], a[0][:, h:2 * h]])), np.ascontiguousarray(np.stack([b[1][:h], a[1][h:2 * h]])))
    for k, op in enumerate(("mul", "add", "sub")):
        hl, hs = O.empty_complex(pb, h)
        assert host.host_elementwise_complex(pb, k, h, P(ca[0], u64p), P(ca[1], i64p), P(cb[0], u64p), P(cb[1], i64p),
                                             P(hl, u64p), P(hs, i64p)) == 0
        assert O.mismatches(O.elementwise_complex(pb, op, ca, cb), (hl, hs), ignore_zero_sign=False) == 0, op


@pytest.mark.parametrize("pb", [32, 64, 96, 256, 1024])
def test_host_double_conversions(host, pb):
    """from_double / to_double (cu_freal.cuh:141-165): exact round trip when
    pb >= 53, RNDN to pb bits otherwise; the known answer of
    demos/sample_fixed.cu:64-66 is (3.625, -0.5) at every precision."""
    e = ctypes.c_int(0)
    rng = np.random.default_rng(pb)
    for d in list(rng.standard_normal(200) * 10.0 ** rng.integers(-30, 30, 200)) + [1.0, -0.75, 3.625, 2.0 ** -1000]:
        r = host.host_roundtrip_double(pb, float(d), ctypes.byref(e))
        if pb >= 64:
            assert r == d
        else:
            m, ex = np.frexp(d)
            q = float(np.round(np.ldexp(m, 32)))       # RNDN to 32 bits (half-even)
            assert r == np.ldexp(q, ex - 32)
    assert host.host_roundtrip_double(pb, 0.0, ctypes.byref(e)) == 0.0
