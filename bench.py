"""Headline benchmark: multiple-precision ops/s of the fixed-precision hot path
on one MI355X.

Workload (the configuration the metric is quoted on): complex AXPY
z = a*x + y in cu_fcomplex<1024>, n = 2^18, Re/Im exponents uniform in
[-16, 16] (seeded synthetic values, full 1024-bit significands), MPC_RNDNN per
operation, bit-exact.  One step = one AXPY over the whole array with the
inputs resident in HBM.  One element = 2 MP ops (one complex multiply, one
complex add), so value = 2 * n * steps / time.

Also printed in the same JSON line: the roofline of the dominant kernel
(algorithmic bytes 816 B/element and word products 4096/element over the
kernel's average launch time, measured with HIP events on the launch stream),
and the CPU baseline: the reference's own cu_fcomplex<1024> host code
(oracle/_ref, compiled from the reference header) over all host cores with
OpenMP, on the same inputs.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--extra]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E, published
PB, N_ELEM = 1024, 1 << 18
BYTES_PER_ELEM = 3 * 2 * (PB // 8 + 8)        # x, y, z; Re+Im; limbs + sign/exp word = 816
WORD_PRODUCTS_PER_ELEM = 4 * (PB // 32) ** 2   # four exact 1024x1024-bit products = 4096


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--extra", action="store_true", help="also time the other configurations")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the path's multi-GPU sharding is described in DESIGN.md but not run", file=sys.stderr)
        sys.exit(2)

    import torch
    import mpc_cuda_amd as M

    torch.cuda.set_device(0)
    a = M.random_complex(PB, 1, 201, -16, 16).cuda()
    x = M.random_complex(PB, N_ELEM, 202, -16, 16).cuda()
    y = M.random_complex(PB, N_ELEM, 203, -16, 16).cuda()
    z = M.empty(PB, N_ELEM, True)
    for _ in range(args.warmup):
        M.axpy(a, x, y, out=z)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        M.axpy(a, x, y, out=z)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps           # HIP events around the launches
    ms_per_step = wall * 1e3 / args.steps
    value = 2.0 * N_ELEM / (ms_per_step * 1e-3)

    achieved_gbs = BYTES_PER_ELEM * N_ELEM / (kern_ms * 1e-3) / 1e9
    wp_rate = WORD_PRODUCTS_PER_ELEM * N_ELEM / (kern_ms * 1e-3)
    rec = {
        "metric": "MP ops/s @256/1024-bit: AXPY, GEMM (MFMA exact-dot vs per-op), 0-ULP",
        "value": value,
        "unit": "MP ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 words (integer limb arithmetic, 1024-bit significands)",
        "data": "synthetic (seeded xorshift64 significands, exponents in [-16,16])",
        "config": {"workload": "complex AXPY z=a*x+y, cu_fcomplex<1024>, n=2^18, MPC_RNDNN per op",
                   "pb": PB, "n": N_ELEM, "mp_ops_per_element": 2},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_axpy_cplx<1024>", "kernel_ms": kern_ms,
                     "algorithmic_bytes_per_launch": BYTES_PER_ELEM * N_ELEM,
                     "word_products_per_s": wp_rate},
    }

    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(a, x, y, z)

    if args.extra:
        rec["extra"] = extra_configs(args)

    print(json.dumps(rec))


def cpu_baseline(a, x, y, z):
    """The reference's cu_fcomplex<1024> AXPY compiled from its own header
    (oracle/_ref), OpenMP over all host cores, on the same inputs as the GPU
    run; its output is also checked against the GPU result."""
    import numpy as np
    from oracle import oracle as O
    ap, xp, yp = a.pair(), x.pair(), y.pair()
    kind = "reference" if O.ref_available() else "port"
    impl = "ref" if kind == "reference" else "oracle"
    cores = O.ref().ref_num_threads() if kind == "reference" else O.lib().or_num_threads()
    reps, t_total = 0, 0.0
    out = None
    while t_total < 8.0 and reps < 50:
        t0 = time.perf_counter()
        out = O.axpy_complex(PB, ap, xp, yp, impl=impl)
        t_total += time.perf_counter() - t0
        reps += 1
    rate = 2.0 * N_ELEM * reps / t_total
    bad = O.mismatches(z.pair(), out, ignore_zero_sign=False)
    return {"value": rate, "unit": "MP ops/s", "cores": int(cores), "kind": kind,
            "sample": f"full workload (n=2^18 complex AXPY at pb=1024) x {reps} repetitions, "
                      f"{t_total:.1f} s wall on {cores} threads",
            "gpu_vs_cpu_mismatches": int(bad)}


def extra_configs(args):
    import torch
    import mpc_cuda_amd as M
    out = {}

    def timed(fn, steps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    for pb in (64, 128, 256, 512, 2048):
        n = 1 << 20
        a = M.random_real(pb, 1, 1, -8, 8).cuda()
        x = M.random_real(pb, n, 2, -8, 8).cuda()
        y = M.random_real(pb, n, 3, -8, 8).cuda()
        o = M.empty(pb, n)
        ms = timed(lambda: M.axpy(a, x, y, out=o), 10)
        out[f"axpy_real_pb{pb}"] = {"ms": ms, "MP_ops_per_s": 2 * n / (ms * 1e-3),
                                    "GB_per_s": 3 * (pb // 8 + 8) * n / (ms * 1e-3) / 1e9,
                                    "word_products_per_s": (pb // 32) ** 2 * n / (ms * 1e-3)}
    pb, m, k = 512, 4096, 4096
    A = M.random_real(pb, m * k, 301, -4, 4).cuda()
    xv = M.random_real(pb, k, 302, -4, 4).cuda()
    ms = timed(lambda: M.gemv(A, xv, m, k), 2)
    out["gemv_pb512_4096"] = {"ms": ms, "MP_ops_per_s": 2 * m * k / (ms * 1e-3),
                              "word_products_per_s": m * k * (pb // 32) ** 2 / (ms * 1e-3)}
    pb, n = 256, 1024
    A = M.random_real(pb, n * n, 401, -4, 4).cuda()
    B = M.random_real(pb, n * n, 402, -4, 4).cuda()
    ms = timed(lambda: M.gemm(A, B, n, n, n, "perop"), 1)
    out["gemm_perop_pb256_1024"] = {"ms": ms, "MP_ops_per_s": 2 * n ** 3 / (ms * 1e-3)}
    return out


if __name__ == "__main__":
    main()
