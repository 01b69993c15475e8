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
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The batched execution's hipGraph runs one branch per (precision, real/complex)
# class; the HIP runtime maps concurrent branches onto its hardware queues (4
# by default), so allow enough queues for the classes to run side by side.
# Read by the runtime at initialisation, hence before anything imports torch.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "12")

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
    ap.add_argument("--driver", choices=["axpy", "exact", "gemv"], default=None,
                    help="only run a few launches of one kernel (for rocprofv3 --pmc): the headline AXPY "
                         "plus calibration copies, the exact GEMM at 1024^3, or the GEMV at 4096^2")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the path's sharding over several GPUs is designed but not run", file=sys.stderr)
        sys.exit(2)

    import torch
    import mpc_cuda_amd as M

    torch.cuda.set_device(0)
    if args.driver:
        profile_driver(args.driver)
        return
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
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_axpy_cplx_pipe<1024>", "kernel_ms": kern_ms,
                     "algorithmic_bytes_per_launch": BYTES_PER_ELEM * N_ELEM,
                     # rate in 32x32-bit word products of the four exact schoolbook
                     # products per element (the kernel computes short products)
                     "equiv_word_products_per_s": wp_rate},
    }

    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(a, x, y, z)

    rec["side_metrics"] = side_metrics()

    if args.extra:
        rec["extra"] = extra_configs(args)

    print(json.dumps(rec))


def profile_driver(which):
    """Launches for the rocprofv3 --pmc passes summarised under profiles/:
    "axpy": three headline AXPYs on device-generated inputs, then three 512 MiB
    device copies whose FETCH_SIZE / WRITE_SIZE calibrate the counters;
    "exact": three exact GEMMs pb 256 1024^3 (exponents in [-4, 4]);
    "gemv": three per-op GEMVs pb 512 4096^2."""
    import torch
    import mpc_cuda_amd as M
    if which == "axpy":
        a = M.device_random(PB, 1, 1, -16, 16, cplx=True)
        x = M.device_random(PB, N_ELEM, 2, -16, 16, cplx=True)
        y = M.device_random(PB, N_ELEM, 3, -16, 16, cplx=True)
        for _ in range(3):
            M.axpy(a, x, y)
        src = torch.empty(1 << 29, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
    elif which == "exact":
        d = 1024
        A = M.device_random(256, d * d, 401, -4, 4)
        B = M.device_random(256, d * d, 402, -4, 4)
        C = M.empty(256, d * d)
        for _ in range(3):
            M.gemm(A, B, d, d, d, "exact", out=C)
    else:
        m = k = 4096
        A = M.device_random(512, m * k, 301, -4, 4)
        x = M.device_random(512, k, 302, -4, 4)
        for _ in range(3):
            M.gemv(A, x, m, k)
    torch.cuda.synchronize()
    print(f"profile driver {which} done", flush=True)


def measured_traffic():
    """HBM bytes per launch of the headline kernel at this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated against a
    copy of known size): profiles/axpy_cplx_pb1024_traffic.md."""
    import re
    try:
        with open(os.path.join(ROOT, "profiles", "axpy_cplx_pb1024_traffic.md")) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", f.read(), re.M)
        return int(m.group(1)) if m else None
    except OSError:
        return None


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


def _timed(fn, steps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def side_metrics():
    """The other halves of the metric, measured in the same run: real AXPY at
    256 bits (n = 2^20) and the 1024^3 GEMM at 256 bits, per-op vs MFMA
    exact-dot (exponents in [-4, 4])."""
    import mpc_cuda_amd as M
    out = {}
    pb, n = 256, 1 << 20
    a = M.device_random(pb, 1, 101, -8, 8)
    x = M.device_random(pb, n, 102, -8, 8)
    y = M.device_random(pb, n, 103, -8, 8)
    o = M.empty(pb, n)
    ms = _timed(lambda: M.axpy(a, x, y, out=o), 20)
    out["axpy_real_pb256_n2^20"] = {"ms": ms, "MP_ops_per_s": 2 * n / (ms * 1e-3),
                                    "GB_per_s": 120 * n / (ms * 1e-3) / 1e9, "hbm_frac": 120 * n / (ms * 1e-3) / 8e12}
    pb, d = 256, 1024
    A = M.device_random(pb, d * d, 401, -4, 4)
    B = M.device_random(pb, d * d, 402, -4, 4)
    C = M.empty(pb, d * d)
    ms_p = _timed(lambda: M.gemm(A, B, d, d, d, "perop", out=C), 1)
    ms_e = _timed(lambda: M.gemm(A, B, d, d, d, "exact", out=C), 2)
    out["gemm_pb256_1024^3"] = {"perop_ms": ms_p, "exact_ms": ms_e, "exact_speedup": ms_p / ms_e,
                                "perop_MP_ops_per_s": 2 * d ** 3 / (ms_p * 1e-3),
                                "exact_MP_ops_per_s": 2 * d ** 3 / (ms_e * 1e-3)}
    return out


def extra_configs(args):
    """Every other configuration, on GPU-generated inputs: real AXPY over the
    precision sweep (n = 2^20), GEMV 4096^2 at 512 bits, the exact GEMM with
    exponent spread [-80, 80], the complex GEMM 1024^3 at 256 bits (per-op and
    exact), division and square root at 256 and 1024 bits (n = 2^18), the
    run-time precision AXPY at 1024 and 1000 bits (n = 16384), and the
    10,000-instance batched workload."""
    import mpc_cuda_amd as M
    out = {}
    for pb in (64, 128, 512, 1024, 2048):
        n = 1 << 20
        a = M.device_random(pb, 1, 1, -8, 8)
        x = M.device_random(pb, n, 2, -8, 8)
        y = M.device_random(pb, n, 3, -8, 8)
        o = M.empty(pb, n)
        ms = _timed(lambda: M.axpy(a, x, y, out=o), 10)
        out[f"axpy_real_pb{pb}"] = {"ms": ms, "MP_ops_per_s": 2 * n / (ms * 1e-3),
                                    "GB_per_s": 3 * (pb // 8 + 8) * n / (ms * 1e-3) / 1e9,
                                    "word_products_per_s": (pb // 32) ** 2 * n / (ms * 1e-3)}
        del a, x, y, o
    pb, m, k = 512, 4096, 4096
    A = M.device_random(pb, m * k, 301, -4, 4)
    xv = M.device_random(pb, k, 302, -4, 4)
    ms = _timed(lambda: M.gemv(A, xv, m, k), 2)
    out["gemv_pb512_4096"] = {"ms": ms, "MP_ops_per_s": 2 * m * k / (ms * 1e-3),
                              "word_products_per_s": m * k * (pb // 32) ** 2 / (ms * 1e-3)}
    del A, xv
    pb, d = 256, 1024
    A = M.device_random(pb, d * d, 501, -80, 80)
    B = M.device_random(pb, d * d, 502, -80, 80)
    ms = _timed(lambda: M.gemm(A, B, d, d, d, "exact"), 2)
    out["gemm_exact_pb256_1024^3_spread80"] = {"ms": ms, "MP_ops_per_s": 2 * d ** 3 / (ms * 1e-3)}
    del A, B
    # complex GEMM pb 256, 1024^3: per-op (MPC_RNDNN per operation) vs exact
    # per component (one rounding); one complex multiply-add = 2 MP ops
    A = M.device_random(pb, d * d, 511, -4, 4, cplx=True)
    B = M.device_random(pb, d * d, 512, -4, 4, cplx=True)
    ms_p = _timed(lambda: M.gemm(A, B, d, d, d, "perop"), 1)
    ms_e = _timed(lambda: M.gemm(A, B, d, d, d, "exact"), 2)
    out["cgemm_pb256_1024^3"] = {"perop_ms": ms_p, "exact_ms": ms_e, "exact_speedup": ms_p / ms_e,
                                 "exact_MP_ops_per_s": 2 * d ** 3 / (ms_e * 1e-3)}
    del A, B
    # division and square root (cu_fdiv / cu_sqrt: Newton at pb + 128 bits),
    # n = 2^18 values, results per second
    for pb in (256, 1024):
        n = 1 << 18
        a = M.device_random(pb, n, 601, -30, 30)
        b = M.device_random(pb, n, 602, -30, 30)
        o = M.empty(pb, n)
        ms_d = _timed(lambda: M.elementwise("div", a, b, out=o), 5)
        ms_s = _timed(lambda: M.sqrt(a, out=o), 5)
        out[f"div_sqrt_pb{pb}_n2^18"] = {"div_ms": ms_d, "divisions_per_s": n / (ms_d * 1e-3),
                                         "sqrt_ms": ms_s, "square_roots_per_s": n / (ms_s * 1e-3)}
        del a, b, o
    # run-time precision AXPY (the reference's MPFR / MPC device demos,
    # N = 16384): at 1024 bits and at 1000 bits (not a compiled precision);
    # host-generated inputs
    for p in (1024, 1000):
        n = 16384
        a = M.random_real(p, 1, 701, -8, 8).cuda()
        x = M.random_real(p, n, 702, -8, 8).cuda()
        y = M.random_real(p, n, 703, -8, 8).cuda()
        ms_r = _timed(lambda: M.rt_axpy(a, x, y), 10)
        ca = M.random_complex(p, 1, 704, -8, 8).cuda()
        cx = M.random_complex(p, n, 705, -8, 8).cuda()
        cy = M.random_complex(p, n, 706, -8, 8).cuda()
        ms_c = _timed(lambda: M.rt_axpy(ca, cx, cy), 10)
        out[f"rt_axpy_p{p}_n16384"] = {"real_ms": ms_r, "complex_ms": ms_c,
                                       "real_MP_ops_per_s": 2 * n / (ms_r * 1e-3),
                                       "complex_MP_ops_per_s": 2 * n / (ms_c * 1e-3)}
        del a, x, y, ca, cx, cy
    out["batch_10000"] = batch_config(args)
    return out


def batch_config(args, n_tasks=10000, seed=5, check=50):
    """10,000 instances: AXPY / dot / GEMV at 40/40/20 %, half complex, lengths
    log-uniform in [16, 65536], GEMV rows log-uniform in [16, 2048] with the
    matrix capped at 2^18 values (columns reduced), precision uniform in
    {64, 128, 256, 512, 1024}, exponents in [-8, 8]; one graph launch.  Fixed
    mapping vs tuned mapping; `check` sampled instances are verified against
    the oracle."""
    import math
    import random
    import torch
    import mpc_cuda_amd as M
    from mpc_cuda_amd import batch as B
    rng = random.Random(seed)
    logu = lambda a, b: int(round(2 ** rng.uniform(math.log2(a), math.log2(b))))
    tasks, keep, specs = [], [], []
    mp_ops, wp = 0, 0
    for t in range(n_tasks):
        kind = rng.choices([B.AXPY, B.DOT, B.GEMV], [0.4, 0.4, 0.2])[0]
        pb = rng.choice([64, 128, 256, 512, 1024])
        cplx = rng.random() < 0.5
        L = logu(16, 65536)
        R = 1
        if kind == B.GEMV:
            R = logu(16, 2048)
            L = max(1, min(L, (1 << 18) // R))
        g = lambda cnt, s: M.device_random(pb, cnt, s, -8, 8, cplx)
        sd = 7919 * t + 13
        if kind == B.AXPY:
            arrs = (g(1, sd), g(L, sd + 1), g(L, sd + 2), M.empty(pb, L, cplx))
            tasks.append(B.axpy(*arrs))
            elems = L
        elif kind == B.DOT:
            arrs = (g(L, sd), g(L, sd + 1), M.empty(pb, 1, cplx))
            tasks.append(B.dot(*arrs))
            elems = L
        else:
            arrs = (g(R * L, sd), g(L, sd + 1), M.empty(pb, R, cplx))
            tasks.append(B.gemv(arrs[0], arrs[1], R, L, arrs[2]))
            elems = R * L
        specs.append((kind, pb, cplx, R, L, arrs))
        mp_ops += 2 * elems
        wp += elems * (pb // 32) ** 2 * (4 if cplx else 1)
    torch.cuda.synchronize()
    res = {"instances": n_tasks, "mp_ops": mp_ops, "word_products": wp}
    for label, tuning in (("fixed", False), ("tuned", True)):
        bt = B.Batch(tasks, tuning=tuning,
                     cache=os.path.join(tempfile.gettempdir(), "mpgfx_batch_tune.json") if tuning else None)
        ms = _timed(bt.run, 3)
        res[label] = {"ms": ms, "instances_per_s": n_tasks / (ms * 1e-3), "MP_ops_per_s": mp_ops / (ms * 1e-3),
                      "word_products_per_s": wp / (ms * 1e-3), "scratch_GB": bt.info()["scratch_bytes"] / 1e9}
        bt.close()
    if check:
        from oracle import oracle as O
        bad = 0
        for i in sorted(rng.sample(range(n_tasks), check)):
            kind, pb, cplx, R, L, arrs = specs[i]
            if kind == B.AXPY:
                ref = (O.axpy_complex if cplx else O.axpy_real)(pb, arrs[0].pair(), arrs[1].pair(), arrs[2].pair())
                got = arrs[3]
            elif kind == B.DOT:
                ref = O.gemm(pb, 1, 1, L, arrs[0].pair(), arrs[1].pair(), "cperop" if cplx else "perop")
                got = arrs[2]
            else:
                ref = O.gemm(pb, R, 1, L, arrs[0].pair(), arrs[1].pair(), "cperop" if cplx else "perop")
                got = arrs[2]
            bad += O.mismatches(got.pair(), ref, ignore_zero_sign=False)
        res["sampled_instances_checked"] = check
        res["sampled_mismatches"] = bad
    return res


if __name__ == "__main__":
    main()
