"""Seeded synthetic MP inputs, generated on the host with numpy.

Each value i draws from its own xorshift64 (13, 7, 17) stream -- the generator
of the reference's tests (tools/test_freal.cu:15-27) -- seeded with
splitmix64(seed, i), so any slice of a large array is reproducible on its own
and generation vectorises over values.  Per value: one word for the sign, N
limb words (top bit forced, low 64N-pb bits cleared), one word for the exponent,
uniform in [emin, emax] as ``xs % (emax-emin+1) + emin``.
"""
import numpy as np

from .arrays import MPArray, nlimbs

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def _xs(s):
    s ^= (s << np.uint64(13)) & _M64
    s ^= s >> np.uint64(7)
    s ^= (s << np.uint64(17)) & _M64
    return s


def random_real(pb, n, seed, emin=-8, emax=8, start=0):
    """n random nonzero values at precision pb (host MPArray)."""
    N = nlimbs(pb)
    SB = 64 * N - pb
    with np.errstate(over="ignore"):
        idx = np.arange(start, start + n, dtype=np.uint64)
        s = _splitmix64(idx ^ _splitmix64(np.uint64(seed)))
        s[s == 0] = np.uint64(1)
        s = _xs(s)
        neg = (s & np.uint64(1)).astype(np.int64)
        limbs = np.empty((N, n), np.uint64)
        for l in range(N):
            s = _xs(s)
            limbs[l] = s
        limbs[N - 1] |= np.uint64(1 << 63)
        if SB:
            limbs[0] &= np.uint64(~((1 << SB) - 1) & 0xFFFFFFFFFFFFFFFF)
        s = _xs(s)
        span = np.uint64(emax - emin + 1)
        e = (s % span).astype(np.int64) + emin
    se = (e << 1) | neg
    return MPArray(pb, limbs, se, False)


def random_complex(pb, n, seed, emin=-8, emax=8):
    """n random complex values with independent real and imaginary parts."""
    re = random_real(pb, n, seed, emin, emax)
    im = random_real(pb, n, seed ^ 0x5DEECE66D, emin, emax)
    return MPArray(pb, np.stack([re.limbs, im.limbs]), np.stack([re.se, im.se]), True)


def random_array(pb, n, seed, emin=-8, emax=8, cplx=False):
    return (random_complex if cplx else random_real)(pb, n, seed, emin, emax)


def device_random(pb, n, seed, emin=-8, emax=8, cplx=False, device="cuda"):
    """The same values as random_real / random_complex, generated on the GPU
    by mpgfx_fill_random."""
    import ctypes
    import torch
    from . import _lib
    from .arrays import empty
    out = empty(pb, n, cplx, device=device)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.lib()
    parts = [(out.limbs, out.se, seed)] if not cplx else [
        (out.limbs[0], out.se[0], seed), (out.limbs[1], out.se[1], seed ^ 0x5DEECE66D)]
    for l, se, sd in parts:
        _lib.check("mpgfx_fill_random", L.mpgfx_fill_random(pb, n, 0, sd & 0xFFFFFFFFFFFFFFFF, emin, emax,
                                                            l.data_ptr(), se.data_ptr(), s))
    return out
