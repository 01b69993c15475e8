"""Benchmark of the fixed-precision hot path on one MI355X.

Headline (the configuration the metric is quoted on): complex AXPY
z = a*x + y in cu_fcomplex<1024>, n = 2^18, Re/Im exponents uniform in
[-16, 16] (seeded synthetic values, full 1024-bit significands), MPC_RNDNN per
operation, bit-exact.  One step = one AXPY over the whole array with the
inputs resident in HBM.  One element = 2 MP ops (one complex multiply, one
complex add), so value = 2 * n * steps / time.

The same JSON line carries, under "configs", every other configuration of the
metric, each measured in this run with its own roofline and CPU baseline:
real AXPY (pb 256, n = 2^20, and the sweep over pb 64/128/512/2048), the
per-op GEMV (pb 512, 4096^2), the per-op and exact GEMM (pb 256, 1024^3,
exponents in [-4, 4] and [-80, 80]), the 10,000-instance batched mix (fixed
and tuned mapping), and the rows beside the path (complex GEMM, division and
square root, run-time precision AXPY).

Rooflines: algorithmic bytes B(pb) = pb/8 + 8 per real value against the
8 TB/s HBM peak; word products (pb/32)^2 per exact real product against the
VALU integer multiply-add peak (one v_mad_u64_u32 per lane per clock:
256 CUs x 64 x 2.4 GHz = 3.93e13/s); int8 operations S^2 2MNK (S = 32
slices) against the 5.0 POP/s dense INT8 MFMA peak.  Kernel times are HIP
events on the launch stream.

CPU baselines: system MPFR / MPC in the kernels' operation order, OpenMP over
the host's cores (oracle/libcpu_mpfr.so, the contender of
demos/bench3_cpu.cpp), on the same inputs, on a bounded sample stated in each
record; its results are compared with the GPU's.  The headline also times the
reference's own cu_fcomplex<1024> header code (oracle/_ref) when it is built.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import math
import os
import platform
import random
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
VALU_WP_PEAK = 256 * 64 * 2.4e9  # v_mad_u64_u32 per lane per clock, 256 CUs
INT8_PEAK = 5.0e15               # dense INT8 MFMA, published
PB, N_ELEM = 1024, 1 << 18
METRIC = "MP ops/s @256/1024-bit: AXPY, GEMM (MFMA exact-dot vs per-op), 0-ULP"


def B(pb):
    return pb // 8 + 8


def WP(pb):
    return (pb // 32) ** 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-configs", action="store_true", help="headline only")
    ap.add_argument("--driver", choices=["axpy", "axpyr", "exact", "gemv", "gemvc", "gemm"], default=None,
                    help="only run a few launches of one kernel (for rocprofv3 --pmc): the headline AXPY "
                         "plus calibration copies, the real AXPY pb 2048 n = 2^20, the exact GEMM at 1024^3, "
                         "the GEMV at 4096^2 (real or complex), or the per-op GEMM at 1024^3")
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

    bytes_launch = 3 * 2 * B(PB) * N_ELEM
    achieved_gbs = bytes_launch / (kern_ms * 1e-3) / 1e9
    wp_rate = 4 * WP(PB) * N_ELEM / (kern_ms * 1e-3)
    rec = {
        "metric": METRIC,
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
                     "kernel": "k_axpy_cplx_mfma<1024>", "kernel_ms": kern_ms,
                     "algorithmic_bytes_per_launch": bytes_launch,
                     # 32x32-bit word products of the four exact schoolbook products per
                     # element, against the VALU peak (an equivalent rate: the kernel forms
                     # the products' top digits as int8 digit products on MFMA)
                     "equiv_word_products_per_s": wp_rate, "valu_frac": wp_rate / VALU_WP_PEAK},
    }
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_headline(a, x, y, z)
    if not args.no_configs:
        rec["configs"] = configs(not args.no_cpu)
    print(json.dumps(rec))


# ----------------------------------------------------------------- helpers

def _timed(fn, steps, warm=1):
    """average milliseconds of fn over `steps` launches, HIP events on the
    current (launch) stream"""
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def host_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def _cpu_common():
    from oracle import oracle as O
    return {"unit": "MP ops/s", "cores": int(O.cpu().cpu_num_threads()), "kind": "port",
            "libraries": O.cpu().cpu_versions().decode(), "cpu_model": host_info()}


def hbm_roof(bytes_, ms):
    gbs = bytes_ / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "algorithmic_bytes_per_launch": bytes_}


def valu_roof(wp, ms):
    r = wp / (ms * 1e-3)
    return {"bound": "valu", "achieved": r, "peak": VALU_WP_PEAK, "unit": "word-products/s",
            "frac": r / VALU_WP_PEAK, "algorithmic_word_products_per_launch": wp}


def cpu_headline(a, x, y, z):
    """MPC (mpc_mul + mpc_add, MPC_RNDNN) on the whole headline workload, and
    the reference's own cu_fcomplex<1024> header code when oracle/_ref is
    built; both checked against the GPU result."""
    from oracle import oracle as O
    ap, xp, yp = a.pair(), x.pair(), y.pair()
    gpu = z.pair()
    out = {}
    if O.cpu_available():
        _, ms1 = O.cpu_axpy(PB, ap, xp, yp, True)
        reps = max(1, min(2000, int(10000.0 / max(ms1, 1e-3))))
        res, ms = O.cpu_axpy(PB, ap, xp, yp, True, reps=reps)
        out = dict(_cpu_common())
        out.update({"value": 2.0 * N_ELEM / (ms * 1e-3),
                    "sample": f"full workload (n=2^18 complex AXPY at pb=1024, mpc_mul + mpc_add) x {reps} "
                              f"repetitions, {ms * reps / 1e3:.1f} s of timed loop",
                    "gpu_vs_cpu_mismatches": O.mismatches(gpu, res, ignore_zero_sign=True)})
    if O.ref_available():
        reps, t_total, res = 0, 0.0, None
        while t_total < 3.0 and reps < 20:
            t0 = time.perf_counter()
            res = O.axpy_complex(PB, ap, xp, yp, impl="ref")
            t_total += time.perf_counter() - t0
            reps += 1
        hdr = {"value": 2.0 * N_ELEM * reps / t_total, "unit": "MP ops/s",
               "cores": int(O.ref().ref_num_threads()), "kind": "reference",
               "sample": f"full workload x {reps} repetitions, the reference's cu_fcomplex<1024> header code "
                         f"(OpenMP), {t_total:.1f} s",
               "gpu_vs_cpu_mismatches": O.mismatches(gpu, res, ignore_zero_sign=False)}
        if out:
            out["reference_header"] = hdr
        else:
            out = hdr
    return out or None


# ----------------------------------------------------------------- configs

def configs(with_cpu):
    from oracle import oracle as O
    with_cpu = with_cpu and O.cpu_available()
    out = {}
    out["axpy_real"] = axpy_real_sweep(with_cpu)
    out["gemv_pb512_4096x4096"] = gemv_config(with_cpu)
    out.update(gemm_configs(with_cpu))
    out["batch_10000"] = batch_config(with_cpu)
    out.update(side_rows())
    return out


def axpy_real_sweep(with_cpu):
    """Real AXPY out = RNDN(RNDN(a*x) + y), n = 2^20: pb 256 (exponents in
    [-8, 8], the metric's 256-bit AXPY) and the sweep pb 64/128/512/2048."""
    import mpc_cuda_amd as M
    from oracle import oracle as O
    res = {}
    for pb in (256, 64, 128, 512, 2048):
        n = 1 << 20
        a = M.device_random(pb, 1, 101 + pb, -8, 8)
        x = M.device_random(pb, n, 102 + pb, -8, 8)
        y = M.device_random(pb, n, 103 + pb, -8, 8)
        o = M.empty(pb, n)
        ms = _timed(lambda: M.axpy(a, x, y, out=o), 20)
        r = {"ms": ms, "value": 2 * n / (ms * 1e-3), "unit": "MP ops/s",
             "roofline": hbm_roof(3 * B(pb) * n, ms), "valu": valu_roof(WP(pb) * n, ms)}
        if with_cpu:
            ns = n if pb <= 512 else 1 << 17
            ap, xs, ys = a.pair(), _slice(x, ns), _slice(y, ns)
            got, cms = O.cpu_axpy(pb, ap, xs, ys, False)
            c = dict(_cpu_common())
            c.update({"value": 2 * ns / (cms * 1e-3),
                      "sample": f"first {ns} elements (mpfr_mul + mpfr_add), {cms:.0f} ms",
                      "gpu_vs_cpu_mismatches": O.mismatches(_slice(o, ns), got, ignore_zero_sign=True)})
            r["cpu_baseline"] = c
        res[f"pb{pb}_n2^20"] = r
        del a, x, y, o
    return res


def _slice(arr, ns):
    import numpy as np
    l, s = arr.limbs[..., :ns], arr.se[..., :ns]
    return np.ascontiguousarray(l.cpu().numpy().view(np.uint64)), np.ascontiguousarray(s.cpu().numpy())


def gemv_config(with_cpu):
    """GEMV y = A x, pb 512, A 4096 x 4096, exponents in [-4, 4], per-op
    k-order fold (tools/linalg_fast.cu:37-51)."""
    import mpc_cuda_amd as M
    from oracle import oracle as O
    pb, m, k = 512, 4096, 4096
    A = M.device_random(pb, m * k, 301, -4, 4)
    xv = M.device_random(pb, k, 302, -4, 4)
    yv = M.empty(pb, m)
    ms = _timed(lambda: M.gemv(A, xv, m, k, out=yv), 3)
    r = {"ms": ms, "value": 2 * m * k / (ms * 1e-3), "unit": "MP ops/s",
         "roofline": valu_roof(WP(pb) * m * k, ms), "hbm": hbm_roof((m * k + 2 * m) * B(pb), ms)}
    if with_cpu:
        rows = 256
        Ah = _slice_rows(A, 0, rows * k)
        got, cms = O.cpu_gemm(pb, rows, 1, k, Ah, xv.pair(), "perop", (0, rows))
        c = dict(_cpu_common())
        c.update({"value": 2 * rows * k / (cms * 1e-3),
                  "sample": f"rows 0..{rows - 1} of 4096 (per-row mpfr_mul + mpfr_add in k order), {cms:.0f} ms",
                  "gpu_vs_cpu_mismatches": O.mismatches(_slice(yv, rows), got, ignore_zero_sign=True)})
        r["cpu_baseline"] = c
    del A
    return r


def _slice_rows(arr, v0, v1):
    import numpy as np
    l, s = arr.limbs[..., v0:v1], arr.se[..., v0:v1]
    return np.ascontiguousarray(l.cpu().numpy().view(np.uint64)), np.ascontiguousarray(s.cpu().numpy())


def gemm_configs(with_cpu):
    """GEMM C = A B, pb 256, 1024^3: per-op RNDN in k order (the reference's
    matmul) and the exact dot on INT8 MFMA (one rounding per element), with
    exponents in [-4, 4], and the exact form again with [-80, 80]."""
    import numpy as np
    import torch
    import mpc_cuda_amd as M
    from oracle import oracle as O
    pb, d = 256, 1024
    out = {}
    A = M.device_random(pb, d * d, 401, -4, 4)
    Bm = M.device_random(pb, d * d, 402, -4, 4)
    C = M.empty(pb, d * d)
    ms_p = _timed(lambda: M.gemm(A, Bm, d, d, d, "perop", out=C), 2)
    Cp = C.pair()
    r = {"ms": ms_p, "value": 2 * d ** 3 / (ms_p * 1e-3), "unit": "MP ops/s",
         "roofline": valu_roof(WP(pb) * d ** 3, ms_p)}
    Ah, Bh = A.pair(), Bm.pair()
    if with_cpu:
        rows = 16
        got, cms = O.cpu_gemm(pb, d, d, d, Ah, Bh, "perop", (0, rows))
        c = dict(_cpu_common())
        c.update({"value": 2 * rows * d * d / (cms * 1e-3),
                  "sample": f"rows 0..{rows - 1} of 1024 (16384 of 1048576 outputs), {cms:.0f} ms; "
                            "whole-GEMM CPU time extrapolates x64",
                  "gpu_vs_cpu_mismatches": O.mismatches(_rows(Cp, 0, rows * d), _rows(got, 0, rows * d),
                                                        ignore_zero_sign=True)})
        r["cpu_baseline"] = c
    out["gemm_pb256_1024^3_perop"] = r
    for tag, lo, hi in (("", -4, 4), ("_spread80", -80, 80)):
        if lo != -4:
            A = M.device_random(pb, d * d, 501, lo, hi)
            Bm = M.device_random(pb, d * d, 502, lo, hi)
            Ah, Bh = A.pair(), Bm.pair()
        # steady state: right after the VALU-bound per-op GEMM the first dozen
        # exact calls run up to 25 % slower while the clocks recover
        ms_e = _timed(lambda: M.gemm(A, Bm, d, d, d, "exact", out=C), 10, warm=15)
        ops = 32 * 32 * 2 * d ** 3
        r = {"ms": ms_e, "value": 2 * d ** 3 / (ms_e * 1e-3), "unit": "MP ops/s",
             "roofline": {"bound": "mfma", "achieved": ops / (ms_e * 1e-3) / 1e12, "peak": INT8_PEAK / 1e12,
                          "unit": "TOP/s (int8)", "frac": ops / (ms_e * 1e-3) / INT8_PEAK,
                          "algorithmic_int8_ops_per_launch": ops,
                          "note": "S^2 2MNK with S = 32 slices of 8 bits per 256-bit significand"}}
        if tag == "":
            r["speedup_vs_perop"] = ms_p / ms_e
        if with_cpu:
            rows = 4
            got, cms = O.cpu_gemm(pb, d, d, d, Ah, Bh, "exact", (0, rows))
            c = dict(_cpu_common())
            c.update({"value": 2 * rows * d * d / (cms * 1e-3),
                      "sample": f"rows 0..{rows - 1} of 1024 (exact products at 512 bits, one mpfr_sum each), "
                                f"{cms:.0f} ms",
                      "gpu_vs_cpu_mismatches": O.mismatches(_rows(C.pair(), 0, rows * d), _rows(got, 0, rows * d),
                                                            ignore_zero_sign=True)})
            r["cpu_baseline"] = c
        out[f"gemm_pb256_1024^3_exact{tag}"] = r
    del A, Bm, C
    torch.cuda.synchronize()
    return out


def _rows(pair, v0, v1):
    import numpy as np
    l, s = pair
    return np.ascontiguousarray(l[..., v0:v1]), np.ascontiguousarray(s[..., v0:v1])


def batch_config(with_cpu, n_tasks=10000, seed=5, check=50):
    """10,000 instances (mpc_cuda_amd.workloads.batch_mix: AXPY / dot / GEMV at
    40/40/20 %, half complex, lengths log-uniform in [16, 65536], GEMV rows in
    [16, 2048], pb in {64, 128, 256, 512, 1024}, exponents in [-8, 8]) in one
    graph launch; fixed vs tuned mapping; `check` sampled instances verified
    against the oracle; the CPU baseline times MPFR / MPC on sampled
    instances and extrapolates by word products."""
    import torch
    import mpc_cuda_amd as M
    from mpc_cuda_amd import batch as Bt
    from mpc_cuda_amd import workloads as W
    from oracle import oracle as O
    insts, tasks = W.batch_mix(n_tasks, seed=seed)
    torch.cuda.synchronize()
    mp_ops = sum(2 * x.elements() for x in insts)
    wp = sum(x.word_products() for x in insts)
    byt = sum((2 if x.cplx else 1) * B(x.pb) * (3 * x.elements() if x.kind == Bt.AXPY else x.elements() + x.L + x.R)
              for x in insts)
    res = {"instances": n_tasks, "mp_ops": mp_ops, "word_products": wp}
    for label, tuning in (("fixed", False), ("tuned", True)):
        bt = Bt.Batch(tasks, tuning=tuning,
                      cache=os.path.join(tempfile.gettempdir(), "mpgfx_batch_tune.json") if tuning else None)
        ms = _timed(bt.run, 3)
        res[label] = {"ms": ms, "instances_per_s": n_tasks / (ms * 1e-3), "MP_ops_per_s": mp_ops / (ms * 1e-3),
                      "roofline": valu_roof(wp, ms), "hbm": hbm_roof(byt, ms),
                      "scratch_GB": bt.info()["scratch_bytes"] / 1e9}
        bt.close()
    rng = random.Random(seed + 1)
    sample = sorted(rng.sample(range(n_tasks), check))
    bad = 0
    for i in sample:
        x = insts[i]
        if x.kind == Bt.AXPY:
            ref = (O.axpy_complex if x.cplx else O.axpy_real)(x.pb, *(v.pair() for v in x.arrays[:3]))
        else:
            ref = O.gemm(x.pb, x.R, 1, x.L, x.arrays[0].pair(), x.arrays[1].pair(), "cperop" if x.cplx else "perop")
        bad += O.mismatches(x.out.pair(), ref, ignore_zero_sign=False)
    res["sampled_instances_checked"] = check
    res["sampled_mismatches"] = bad
    if with_cpu:
        cms, cwp, cbad = 0.0, 0, 0
        for i in sample[:25]:
            x = insts[i]
            if x.kind == Bt.AXPY:
                got, ms = O.cpu_axpy(x.pb, *(v.pair() for v in x.arrays[:3]), x.cplx)
            else:
                got, ms = O.cpu_gemm(x.pb, x.R, 1, x.L, x.arrays[0].pair(), x.arrays[1].pair(),
                                     "cperop" if x.cplx else "perop", (0, x.R))
            cms += ms
            cwp += x.word_products()
            cbad += O.mismatches(x.out.pair(), got, ignore_zero_sign=True)
        c = dict(_cpu_common())
        c.update({"unit": "instances/s", "value": n_tasks / (cms * 1e-3 * wp / cwp),
                  "sample": f"25 sampled instances ({cwp:.3g} word products, {cms:.0f} ms of MPFR / MPC), "
                            f"extrapolated to the batch's {wp:.3g} word products",
                  "gpu_vs_cpu_mismatches": cbad})
        res["cpu_baseline"] = c
    del insts, tasks
    torch.cuda.synchronize()
    return res


def side_rows():
    """Rows beside the path: complex GEMM pb 256 1024^3 (per-op and exact),
    complex GEMV pb 512 4096^2, division and square root (pb 256 and 1024, n = 2^18), run-time precision
    AXPY (1024 and 1000 bits, n = 16384)."""
    import mpc_cuda_amd as M
    out = {}
    pb, d = 256, 1024
    A = M.device_random(pb, d * d, 511, -4, 4, cplx=True)
    Bm = M.device_random(pb, d * d, 512, -4, 4, cplx=True)
    ms_p = _timed(lambda: M.gemm(A, Bm, d, d, d, "perop"), 1)
    ms_e = _timed(lambda: M.gemm(A, Bm, d, d, d, "exact"), 2)
    out["cgemm_pb256_1024^3"] = {"perop_ms": ms_p, "exact_ms": ms_e, "exact_speedup": ms_p / ms_e,
                                 "perop_roofline": valu_roof(4 * WP(pb) * d ** 3, ms_p),
                                 "exact_MP_ops_per_s": 2 * d ** 3 / (ms_e * 1e-3)}
    del A, Bm
    pb, m = 512, 4096
    A = M.device_random(pb, m * m, 521, -4, 4, cplx=True)
    xv = M.device_random(pb, m, 522, -4, 4, cplx=True)
    ms_g = _timed(lambda: M.gemv(A, xv, m, m), 2)
    out["gemv_complex_pb512_4096x4096"] = {"ms": ms_g, "value": 2 * m * m / (ms_g * 1e-3), "unit": "MP ops/s",
                                           "roofline": valu_roof(4 * WP(pb) * m * m, ms_g)}
    del A, xv
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
    return out


# ----------------------------------------------------------------- profiling

def profile_driver(which):
    """Launches for the rocprofv3 passes summarised under profiles/:
    "axpy": three headline AXPYs on device-generated inputs, then three 512 MiB
    device copies whose FETCH_SIZE / WRITE_SIZE calibrate the counters;
    "axpyr": three real AXPYs pb 2048, n = 2^20; "gemvc": two complex GEMVs pb
    512 4096^2;
    "exact": three exact GEMMs pb 256 1024^3 (exponents in [-4, 4]);
    "gemv": three per-op GEMVs pb 512 4096^2; "gemm": two per-op GEMMs pb 256
    1024^3."""
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
    elif which == "axpyr":
        pb, n = 2048, 1 << 20
        a = M.device_random(pb, 1, 101 + pb, -8, 8)
        x = M.device_random(pb, n, 102 + pb, -8, 8)
        y = M.device_random(pb, n, 103 + pb, -8, 8)
        for _ in range(3):
            M.axpy(a, x, y)
    elif which == "gemvc":
        m = k = 4096
        A = M.device_random(512, m * k, 521, -4, 4, cplx=True)
        x = M.device_random(512, k, 522, -4, 4, cplx=True)
        for _ in range(2):
            M.gemv(A, x, m, k)
    elif which == "exact":
        d = 1024
        A = M.device_random(256, d * d, 401, -4, 4)
        Bm = M.device_random(256, d * d, 402, -4, 4)
        C = M.empty(256, d * d)
        for _ in range(3):
            M.gemm(A, Bm, d, d, d, "exact", out=C)
    elif which == "gemm":
        d = 1024
        A = M.device_random(256, d * d, 401, -4, 4)
        Bm = M.device_random(256, d * d, 402, -4, 4)
        C = M.empty(256, d * d)
        for _ in range(2):
            M.gemm(A, Bm, d, d, d, "perop", out=C)
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


if __name__ == "__main__":
    main()
