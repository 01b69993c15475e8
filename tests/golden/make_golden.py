"""Writes the golden fixtures of tests/golden/ (run in a container with
GMP/MPFR/MPC, i.e. where oracle/build/pin_mpfr is built):

    python tests/golden/make_golden.py

Inputs are the structured edge cases of tests/mpcases.py plus values from the
reference's own test generator (xorshift64 13/7/17 with the exponent-spread mix
of tools/test_freal.cu:55) and the package generator; the expected outputs are
computed by GNU MPFR / MPC (mpfr_mul/add/sub RNDN, mpc_mul/add/sub MPC_RNDNN,
the mul+add fold of tools/linalg_fast.cu:71-83, and exact products summed by
mpfr_sum for the exact dot).  Each fixture is a .npz of SoA arrays
(``*_limbs`` uint64, ``*_se`` int64).  MPFR writes IEEE zero signs; tests
compare zeros by value, as the reference's tests do (mpfr_equal_p).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O          # noqa: E402  (MPFR driver)
from mpcases import edge_arrays         # noqa: E402
import mpc_cuda_amd.gen as G            # noqa: E402

PBS = [32, 64, 96, 128, 160, 192, 256, 288, 320, 384, 512, 1024, 1056, 2048]


def cat(a, b):
    return (np.ascontiguousarray(np.concatenate([a[0], b[0]], axis=-1)),
            np.ascontiguousarray(np.concatenate([a[1], b[1]], axis=-1)))


def real_operands(pb, n_ref=32):
    ea, eb, _, _ = edge_arrays(pb, seed=7, n_random=24)
    l, s = O.fill_ref(pb, 2 * n_ref, 0x1111 * (pb // 32), -1)
    ra = (np.ascontiguousarray(l[:, 0::2]), np.ascontiguousarray(s[0::2]))
    rb = (np.ascontiguousarray(l[:, 1::2]), np.ascontiguousarray(s[1::2]))
    return cat((ea.limbs, ea.se), ra), cat((eb.limbs, eb.se), rb)


def cplx(re, im):
    return (np.ascontiguousarray(np.stack([re[0], im[0]])), np.ascontiguousarray(np.stack([re[1], im[1]])))


def main():
    assert O.pin_available(), "oracle/build/pin_mpfr is required (make -C oracle)"
    N = O.nlimbs
    for pb in PBS:
        a, b = real_operands(pb)
        n = a[1].shape[0]
        out = {"a_limbs": a[0], "a_se": a[1], "b_limbs": b[0], "b_se": b[1]}
        for op in ("mul", "add", "sub"):
            r = O.pin(op, pb, [a, b], ((N(pb), n), (n,)), [n])
            out[op + "_limbs"], out[op + "_se"] = r
        # complex: pair real operands as (re, im) with a rotation for variety
        h = n // 2
        ca = cplx((a[0][:, :h], a[1][:h]), (b[0][:, h:2 * h], b[1][h:2 * h]))
        cb = cplx((b[0][:, :h], b[1][:h]), (a[0][:, h:2 * h], a[1][h:2 * h]))
        out.update({"ca_limbs": ca[0], "ca_se": ca[1], "cb_limbs": cb[0], "cb_se": cb[1]})
        for op in ("mul", "add", "sub"):
            r = O.pin("c" + op, pb, [ca, cb], ((2, N(pb), h), (2, h)), [h])
            out["c" + op + "_limbs"], out["c" + op + "_se"] = r
        # AXPY, real and complex
        na = 48
        ax = G.random_real(pb, 1, 11 + pb, -8, 8)
        xx = G.random_real(pb, na, 12 + pb, -8, 8)
        yy = G.random_real(pb, na, 13 + pb, -8, 8)
        r = O.pin("axpy", pb, [(ax.limbs, ax.se), (xx.limbs, xx.se), (yy.limbs, yy.se)],
                  ((N(pb), na), (na,)), [na])
        out.update({"ax_limbs": ax.limbs, "ax_se": ax.se, "xx_limbs": xx.limbs, "xx_se": xx.se,
                    "yy_limbs": yy.limbs, "yy_se": yy.se, "axpy_limbs": r[0], "axpy_se": r[1]})
        cax = G.random_complex(pb, 1, 21 + pb, -16, 16)
        cxx = G.random_complex(pb, na, 22 + pb, -16, 16)
        cyy = G.random_complex(pb, na, 23 + pb, -16, 16)
        r = O.pin("caxpy", pb, [(cax.limbs, cax.se), (cxx.limbs, cxx.se), (cyy.limbs, cyy.se)],
                  ((2, N(pb), na), (2, na)), [na])
        out.update({"cax_limbs": cax.limbs, "cax_se": cax.se, "cxx_limbs": cxx.limbs, "cxx_se": cxx.se,
                    "cyy_limbs": cyy.limbs, "cyy_se": cyy.se, "caxpy_limbs": r[0], "caxpy_se": r[1]})
        # small GEMMs: per-op, exact (two exponent spreads), complex per-op
        m, nn, k = 5, 4, 24
        for tag, (lo, hi) in (("g4", (-4, 4)), ("g80", (-80, 80))):
            A = G.random_real(pb, m * k, 31 + pb + hi, lo, hi)
            B = G.random_real(pb, k * nn, 32 + pb + hi, lo, hi)
            out.update({tag + "A_limbs": A.limbs, tag + "A_se": A.se, tag + "B_limbs": B.limbs, tag + "B_se": B.se})
            for mode in ("perop", "exact"):
                r = O.pin("gemm_" + mode, pb, [(A.limbs, A.se), (B.limbs, B.se)],
                          ((N(pb), m * nn), (m * nn,)), [m, nn, k])
                out[tag + mode + "_limbs"], out[tag + mode + "_se"] = r
        CA = G.random_complex(pb, m * k, 41 + pb, -4, 4)
        CB = G.random_complex(pb, k * nn, 42 + pb, -4, 4)
        r = O.pin("cgemm_perop", pb, [(CA.limbs, CA.se), (CB.limbs, CB.se)],
                  ((2, N(pb), m * nn), (2, m * nn)), [m, nn, k])
        out.update({"cgA_limbs": CA.limbs, "cgA_se": CA.se, "cgB_limbs": CB.limbs, "cgB_se": CB.se,
                    "cgperop_limbs": r[0], "cgperop_se": r[1]})
        r = O.pin("cgemm_exact", pb, [(CA.limbs, CA.se), (CB.limbs, CB.se)],
                  ((2, N(pb), m * nn), (2, m * nn)), [m, nn, k])
        out.update({"cgexact_limbs": r[0], "cgexact_se": r[1]})
        out["dims"] = np.array([m, nn, k], np.int64)
        path = os.path.join(HERE, f"pb{pb}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
