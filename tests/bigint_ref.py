"""Independent exact-arithmetic reference (test infrastructure).

Values are (sign, exp, M) with value = sign * M * 2^(exp - 64N), exactly as the
MP arrays store them.  Every operation is done exactly with Python integers and
rounded once with the textbook round-half-to-even of an exact dyadic rational
to pb significant bits.  It shares no code with the C oracle (oracle/mpfixed.c)
or with MPFR, so agreement pins the oracle's rounding logic independently.
"""
ZERO_EXP = -0x7FFFFFFFFFFFFF


def nlimbs(pb):
    return (pb + 63) // 64


def to_exact(v, pb):
    """(sign, exp, M, zero) -> (numerator, exponent2) with value = num * 2^e2."""
    s, e, M, z = v
    if z:
        return 0, 0
    return s * M, e - 64 * nlimbs(pb)


def round_exact(num, e2, pb):
    """RNDN of num * 2^e2 to pb bits -> (sign, exp, M, zero) in storage form
    (M left-justified in 64N bits, zero -> sign +1)."""
    if num == 0:
        return (1, ZERO_EXP, 0, True)
    s = -1 if num < 0 else 1
    a = -num if num < 0 else num
    bl = a.bit_length()
    if bl > pb:
        sh = bl - pb
        q, r = a >> sh, a & ((1 << sh) - 1)
        half = 1 << (sh - 1)
        if r > half or (r == half and (q & 1)):
            q += 1
        if q.bit_length() > pb:          # carry out to 2^pb
            q >>= 1
            sh += 1
        a, e2 = q, e2 + sh
    # exponent: value in [2^(exp-1), 2^exp)
    exp = e2 + a.bit_length()
    M = a << (64 * nlimbs(pb) - a.bit_length())
    return (s, exp, M, False)


def mul(x, y, pb):
    nx, ex = to_exact(x, pb)
    ny, ey = to_exact(y, pb)
    return round_exact(nx * ny, ex + ey, pb)


def add(x, y, pb):
    nx, ex = to_exact(x, pb)
    ny, ey = to_exact(y, pb)
    if nx == 0:          # 0 + y returns y as is (cu_freal.cuh:283-284)
        return y
    if ny == 0:
        return x
    e = min(ex, ey)
    return round_exact((nx << (ex - e)) + (ny << (ey - e)), e, pb)


def neg(v):
    s, e, M, z = v
    return (-s, e, M, z)


def sub(x, y, pb):
    return add(x, neg(y), pb)


def sum_products(pairs, pb):
    """RNDN(sum a*b) with exact products and exact sum."""
    terms = []
    for a, b in pairs:
        na, ea = to_exact(a, pb)
        nb, eb = to_exact(b, pb)
        if na and nb:
            terms.append((na * nb, ea + eb))
    if not terms:
        return (1, ZERO_EXP, 0, True)
    e = min(t[1] for t in terms)
    return round_exact(sum(n << (ee - e) for n, ee in terms), e, pb)


def cmul(ar, ai, br, bi, pb):
    return (sum_products([(ar, br), (neg(ai), bi)], pb), sum_products([(ar, bi), (ai, br)], pb))


def fold_dot(avals, bvals, pb):
    acc = (1, ZERO_EXP, 0, True)
    for a, b in zip(avals, bvals):
        acc = add(acc, mul(a, b, pb), pb)
    return acc
