"""MP arrays: structure-of-arrays storage of cu_freal<PB> / cu_fcomplex<PB>
values, on the host (numpy) or in device memory (torch tensors used purely as
HIP allocations).

Layout (include/mpgfx.h): a real array of n values at precision pb has
``limbs`` of shape (N, n) -- limb l of value i at limbs[l, i], N = ceil(pb/64),
uint64, MPFR-compatible left-justified significand -- and ``se`` of shape (n,)
with se = exp*2 + (sign < 0).  A complex array stores the real array of its
real parts followed by the real array of its imaginary parts: limbs (2, N, n),
se (2, n).
"""
import numpy as np

ZERO_EXP = -0x7FFFFFFFFFFFFF


def nlimbs(pb):
    return (pb + 63) // 64


class MPArray:
    """n values at precision pb; complex if cplx.  limbs/se are numpy arrays
    (host) or int64 torch tensors (device; uint64 bit patterns viewed as
    int64)."""

    def __init__(self, pb, limbs, se, cplx=False):
        self.pb, self.limbs, self.se, self.cplx = pb, limbs, se, cplx
        N = nlimbs(pb)
        n = se.shape[-1]
        want_l = (2, N, n) if cplx else (N, n)
        want_s = (2, n) if cplx else (n,)
        if tuple(limbs.shape) != want_l or tuple(se.shape) != want_s:
            raise ValueError(f"bad MP array shapes {tuple(limbs.shape)} {tuple(se.shape)}, want {want_l} {want_s}")

    @property
    def n(self):
        return self.se.shape[-1]

    @property
    def on_device(self):
        return not isinstance(self.se, np.ndarray)

    # ---- host/device transfer -------------------------------------------
    def cuda(self, device="cuda"):
        import torch
        if self.on_device:
            return self
        l = torch.from_numpy(np.ascontiguousarray(self.limbs).view(np.int64)).to(device)
        s = torch.from_numpy(np.ascontiguousarray(self.se)).to(device)
        return MPArray(self.pb, l, s, self.cplx)

    def numpy(self):
        if not self.on_device:
            return self
        l = self.limbs.cpu().numpy().view(np.uint64)
        s = self.se.cpu().numpy()
        return MPArray(self.pb, l, s, self.cplx)

    def ptrs(self):
        """(limbs pointer, se pointer) of a device array, for the C ABI."""
        assert self.on_device and self.limbs.is_contiguous() and self.se.is_contiguous()
        return self.limbs.data_ptr(), self.se.data_ptr()

    def pair(self):
        """(limbs, se) numpy pair (the oracle's calling convention)."""
        h = self.numpy()
        return np.ascontiguousarray(h.limbs), np.ascontiguousarray(h.se)

    # ---- views ----------------------------------------------------------
    def real_part(self):
        assert self.cplx
        return MPArray(self.pb, self.limbs[0], self.se[0], False)

    def imag_part(self):
        assert self.cplx
        return MPArray(self.pb, self.limbs[1], self.se[1], False)


def empty(pb, n, cplx=False, device="cuda"):
    N = nlimbs(pb)
    if device is None or device == "cpu":
        l = np.zeros((2, N, n) if cplx else (N, n), np.uint64)
        s = np.zeros((2, n) if cplx else (n,), np.int64)
    else:
        import torch
        l = torch.empty((2, N, n) if cplx else (N, n), dtype=torch.int64, device=device)
        s = torch.empty((2, n) if cplx else (n,), dtype=torch.int64, device=device)
    return MPArray(pb, l, s, cplx)


def equal(x, y, ignore_zero_sign=False):
    """Bit-exact comparison of two MP arrays; returns the number of values that
    differ.  ignore_zero_sign treats zeros of either sign field as equal (the
    comparison mpfr_equal_p makes)."""
    xh, yh = x.numpy(), y.numpy()
    xl, xs = xh.limbs, xh.se.copy()
    yl, ys = yh.limbs, yh.se.copy()
    if ignore_zero_sign:
        for s in (xs, ys):
            s[(s >> 1) == ZERO_EXP] = ZERO_EXP * 2
    bad = (xs != ys) | np.any(xl != yl, axis=-2)
    return int(np.count_nonzero(bad))


def decode(limbs, se, pb):
    """Python (sign, exponent, significand-int, is_zero) of one real value."""
    N = nlimbs(pb)
    M = 0
    for l in range(N - 1, -1, -1):
        M = (M << 64) | int(limbs[l])
    w = int(se)
    e = w >> 1
    return (-1 if w & 1 else 1), e, M, e == ZERO_EXP
