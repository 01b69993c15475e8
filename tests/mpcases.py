"""Structured edge-case operands for the RNDN arithmetic (test infrastructure):
ties to even, carry-out to 2^pb, exact and massive cancellation, far operands
(sticky-only and sticky-borrow), equal magnitudes, zeros, all-ones and
single-bit significands -- next to random values.  Returned as host MPArrays so
the same operands feed the HIP library, the C oracle and MPFR."""
import random

import numpy as np

from mpc_cuda_amd.arrays import MPArray, nlimbs, ZERO_EXP


def _pack(vals, pb):
    N = nlimbs(pb)
    n = len(vals)
    limbs = np.zeros((N, n), np.uint64)
    se = np.zeros((n,), np.int64)
    for i, (s, e, M, z) in enumerate(vals):
        if z:
            se[i] = ZERO_EXP * 2 + (1 if s < 0 else 0)
            continue
        for l in range(N):
            limbs[l, i] = (M >> (64 * l)) & 0xFFFFFFFFFFFFFFFF
        se[i] = e * 2 + (1 if s < 0 else 0)
    return MPArray(pb, limbs, se, False)


def _norm(m, pb):
    """significand integer with exactly pb significant bits -> storage M"""
    assert m.bit_length() == pb
    return m << (64 * nlimbs(pb) - pb)


def edge_pairs(pb, seed=1, n_random=64):
    """Lists A, B of (sign, exp, M, zero) tuples, pairwise operands."""
    rng = random.Random(seed * 1000003 + pb)
    top = 1 << (pb - 1)
    ones = (1 << pb) - 1
    A, B = [], []

    def val(m, e, s=1):
        return (s, e, _norm(m, pb), False)

    zero = (1, ZERO_EXP, 0, True)
    rnd = lambda: top | rng.getrandbits(pb - 1)
    cases = [
        (val(ones, 0), val(ones, 0)),                  # mul carry-out, add carry
        (val(ones, 0), val(top, -pb)),                 # add: exact tie ... rounds to even
        (val(ones, 0), val(top, -pb, -1)),             # sub of half ulp
        (val(top, 0), val(top, -pb, -1)),              # sub borrow across the top
        (val(top, 1), val(ones, 0, -1)),               # massive cancellation
        (val(top | 1, 0), val(top, 0, -1)),            # cancellation to one bit
        (val(ones, 5), val(ones, 5, -1)),              # exact cancellation -> +0
        (val(top, 0), val(top, -3 * pb)),              # far: sticky only
        (val(top, 0), val(top, -3 * pb, -1)),          # far: sticky borrow
        (val(top, 0), val(ones, -pb - 1, -1)),         # near-far borrow
        (val(top | 1, 0), val(top | 1, -pb - 1)),      # tie region
        (val(top, 0), val(top, 0)),                    # powers of two
        (val(ones, 3), zero),                          # zero operand
        (zero, val(ones, -3, -1)),
        (zero, zero),
        (val(top + 3, 0), val(top + 5, -2)),
        (val(ones, 10), val(1 << (pb - 1), -10, -1)),
    ]
    # ties for multiplication: (2^(pb-1)+1)^2 lands near halfway for small pb
    cases.append((val(top + 1, 0), val(top + 1, 0)))
    cases.append((val(top + (1 << (pb // 2)), 0), val(top + (1 << (pb // 2)), 0)))
    # ties that only the sticky material decides: a +- (half an ulp of a)(1 + 2^-t)
    # for a of even and odd last bit, the extra bit t places below the half ulp
    # (inside the guard limb, just below it, far below: the sticky borrow of a
    # subtraction and the sticky of an addition); and a carry-out whose
    # dropped bit decides the tie: (2^pb - 1) + 2 (1 + 2^-65) ulps
    for t in sorted({1, 2, 31, 32, 63, 64, 65, 66, 100, pb - 2, pb - 1}):
        if 1 <= t <= pb - 1:
            half = val(top | (1 << (pb - 1 - t)), -pb)
            for a in (val(top | 2, 0), val(top | 3, 0), val(ones - 1, 0), val(ones, 0)):
                cases.append((a, half))
                cases.append((a, (-1, half[1], half[2], half[3])))
    if pb >= 67:
        cases.append((val(ones, 0), val(top | (1 << (pb - 66)), 2 - pb)))
        cases.append((val(ones - 2, 0), val(top | (1 << (pb - 66)), 2 - pb)))
    for a, b in cases:
        A.append(a); B.append(b)
        A.append(b); B.append(a)
    for _ in range(n_random):
        d = rng.choice([0, 1, 2, 31, 32, 33, 63, 64, 65, pb - 1, pb, pb + 1, pb + 2, 2 * pb, rng.randint(0, 3 * pb)])
        ea = rng.randint(-50, 50)
        sa, sb = rng.choice([1, -1]), rng.choice([1, -1])
        ma = rnd()
        mb = rnd() if rng.random() < 0.5 else (ma ^ rng.getrandbits(rng.randint(1, 8)))
        mb |= top
        A.append(val(ma, ea, sa)); B.append(val(mb, ea - d, sb))
    return A, B


def edge_arrays(pb, seed=1, n_random=64):
    A, B = edge_pairs(pb, seed, n_random)
    return _pack(A, pb), _pack(B, pb), A, B


def unpack(arr):
    """host MPArray (real) -> list of (sign, exp, M, zero)"""
    h = arr.numpy()
    out = []
    N = nlimbs(arr.pb)
    for i in range(h.n):
        M = 0
        for l in range(N - 1, -1, -1):
            M = (M << 64) | int(h.limbs[l, i])
        w = int(h.se[i])
        e = w >> 1
        out.append((-1 if w & 1 else 1, e, M, e == ZERO_EXP))
    return out


def pack(vals, pb):
    return _pack(vals, pb)
