#!/usr/bin/env python3
"""Benchmark of the DD hot path on one MI355X.

Default: every workload of the measurement plan in one run, printed as ONE
JSON line. The headline (metric, value, roofline, cpu_baseline) is the
float64x2 GEMM at m = n = k = 4096, entries s*2^U(-20,20)*u with lo per the
config-0 rule, inputs and output resident in HBM:
  ozaki  the INT8-MFMA split-precision GEMM, automatic form (slice count
         chosen on the device, every output certified to 2^-100 (|A||B|)_ij),
         whole call per step;
  plain  the reference-order FP64 GEMM (bit-identical to the reference),
         renorm interval 8.
value = Ozaki DD-GFLOP/s (2 m n k per step) when its measured error on a
sampled block is within 2^-100 (|A||B|) elementwise. The key "configs" holds
the five workloads, each with its own roofline, CPU baseline and accuracy:
  C0  plain GEMM 256^3, config-0 entries (bit-identical to the reference);
  C1  elementwise + - * / sqrt exp log sin erf and sum / dot / norm2 at
      N = 2^26 (the C1 input distributions);
  C2  the headline pair, plus the complex64x2 GEMM at 2048^3;
  C3  Ozaki 8192^3 at exponent spreads 2^+-10 and 2^+-30, plain 8192^3 beside;
  C4  10,000 GEMMs + 50,000 dot products in one persistent grouped launch
      replayed from a HIP graph, per-call launches as the comparison.

  python bench.py [--steps K] [--warmup W] [--workload all|gemm|cgemm|blas|trsm|batched|elementwise|reduce]
                  [--no-cpu] [--quick]

Timing: torch.cuda events recorded on the stream the kernels are launched on
(torch's current stream, whose handle is passed to the C ABI), bracketed by
device synchronisation. CPU baselines run on one host core over a bounded
sample of each workload (stated in each "sample"): the reference's own
compiled code (oracle/_ref: matmuldd_mm, and array loops over its scalar
expdd, logdd, ...) where it has an entry point ("reference"), else the
oracle's C restatement ("port": the Fortran-only reductions).
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
FP64_ISSUE_PEAK = 39.3e12    # FP64 lane-instructions/s: 256 CUs x 64 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_FLOPS_PER_DDMAC = 15.75  # 2 MUL + 2 FMA(x2) + 9 ADD + 6/8 renorm at ri = 8
FP64_INSTR_PER_DDMAC = 13.75  # 2 MUL + 2 FMA + 9 ADD + 6/8 renorm instructions
INT8_PEAK_TOPS = 5000.0      # MI355X dense INT8 MFMA

# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB
# per dispatch), scaled by what the same counters read for a 1 GiB device
# copy: FETCH_SIZE x 2.0 (it reads 524,307 KiB for 1 GiB fetched), WRITE_SIZE
# x 1.0 (1,048,576 KiB). From profiles/hot_kernels_profile_workload_pmc.md
# (bench.py --workload profile). Keyed by (kernel, size, slices).
PMC_TRAFFIC = {
    ("k_oz2_gemm", 4096, 14): (2.0 * 1539677 + 562024) * 1024,
    ("k_gemm_plain", 4096, 0): (2.0 * 3142425 + 262144) * 1024,
}
# the C1 element kernels and reduction passes at N = 2^26, same runs
C1_TRAFFIC = {
    "add": (2.0 * 1048599 + 1048578) * 1024, "sub": (2.0 * 1048580 + 1048578) * 1024,
    "mul": (2.0 * 1048586 + 1048576) * 1024, "div": (2.0 * 1048582 + 1048576) * 1024,
    "sqrt": (2.0 * 524300 + 1048645) * 1024, "exp": (2.0 * 524354 + 1048576) * 1024,
    "log": (2.0 * 524384 + 1048576) * 1024, "sin": (2.0 * 524387 + 1048576) * 1024,
    "erf": (2.0 * 524477 + 1048576) * 1024, "dot": (2.0 * 1048625 + 1024) * 1024,
    "sum": (2.0 * 524334 + 512) * 1024, "norm2": (2.0 * 524371 + 512) * 1024,
}

# FP64 VALU lane-instructions per element of the transcendental array
# kernels on the C1 input distributions: (SQ_INSTS_VALU_ADD_F64 +
# SQ_INSTS_VALU_MUL_F64 + SQ_INSTS_VALU_FMA_F64) x 64 / N from rocprofv3
# --pmc at N = 2^26 (profiles/hot_kernels_profile_workload_pmc.md). With these
# the FP64-issue fraction is instructions/element x elements/s / 39.3e12
# (other FP64 work -- ldexp, rounding, conversions -- is not in these three
# counters, so the fraction is a lower bound of the FP64 pipe's use).
FP64_INSTR_PER_ELEM = {"exp": 319.0, "log": 307.3, "sin": 946.0, "erf": 356.0}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="all",
                   choices=["all", "gemm", "cgemm", "blas", "trsm", "batched", "elementwise", "reduce", "profile",
                            "int8ref", "products"])
    p.add_argument("--retune", action="store_true", help="batched: rerun the tile autotuner first")
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--spread", type=float, default=20.0)
    p.add_argument("--slices", type=int, default=0, help="0 = automatic (certified) form")
    p.add_argument("--form", default="both", choices=["both", "plain", "ozaki"])
    p.add_argument("--err-sample", type=int, default=256, help="error on an N x N block of sampled outputs")
    p.add_argument("--variant", type=int, default=-1)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-cols", type=int, default=128, help="CPU GEMM baseline: all of A times this many columns of B")
    p.add_argument("--quick", action="store_true", help="all: skip C3 and C4 (a shorter smoke of the bench)")
    return p.parse_args()


def timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per step


def host_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def ref_libs():
    """(matmul library, array-loop library) of the compiled reference, or None."""
    from tests import _native as nt  # test infrastructure: paths of the CPU oracle builds
    if not os.path.exists(nt.REF):
        return None, None
    L = ctypes.CDLL(nt.REF)
    L.matmuldd_mm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64] * 4
    loops = os.path.join(os.path.dirname(nt.REF), "librefloop.so")
    R = None
    if os.path.exists(loops):
        R = ctypes.CDLL(loops)
        R.refloop_unary.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        R.refloop_binary.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    return L, R


def cpu_matmul(A, B, cols, ri, reps=1):
    """Time the reference matmul (the oracle's restatement if the reference
    build is absent) on one core over all of A times B[:, :cols]; returns
    (seconds, C block, kind)."""
    from tests import _native as nt
    from multifloats_amd import datagen
    m, k = A.shape[:2]
    Bs = datagen.as_colmajor(B[:, :cols])
    A = datagen.as_colmajor(A)
    C = datagen.as_colmajor(np.zeros((m, cols, 2)))
    L, _ = ref_libs()
    kind, fn = ("reference", L.matmuldd_mm) if L is not None else ("port", nt.oracle().or_matmul_mm)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fn(A.ctypes.data, Bs.ctypes.data, C.ctypes.data, m, k, cols, ri)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, C, kind


def sampled_error(A, B, C, rows, cols):
    from tests import _native as nt  # exact accumulator (test infrastructure)
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    err, mag = nt.gemm_err(A, B, C, rr.ravel(), cc.ravel(), threads=min(16, os.cpu_count() or 1))
    rel = err / mag
    return float(rel.max()), float(rel.mean())


def gemm_roofline_plain(m, n, k, ms, size_key=None, variant=-1):
    achieved = FP64_FLOPS_PER_DDMAC * m * n * k / (ms * 1e-3) / 1e12
    issue = FP64_INSTR_PER_DDMAC * m * n * k / (ms * 1e-3) / FP64_ISSUE_PEAK
    return {"bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "fp64_issue_frac": round(issue, 4),
            "traffic": PMC_TRAFFIC.get(("k_gemm_plain", size_key, 0)) if variant < 0 else None,
            "note": "k_gemm_plain, FP64 VALU: %.2f FP64 flops (%.2f FP64 instructions) per DD-MAC, FMA = 2 flops; "
                    "fp64_issue_frac = instructions issued / (256 CUs x 64 lanes x 2.4 GHz)"
                    % (FP64_FLOPS_PER_DDMAC, FP64_INSTR_PER_DDMAC)}


def ozaki_run(torch, mf, dA, dB, dC, m, k, n, steps, warmup, stream, slices=0, power=False):
    """Time the Ozaki call (automatic form unless slices > 0), then its parts:
    the same call with the slice count it chose (explicit: no certificate)
    and the residue GEMMs alone (k_oz2_gemm, the dominant kernel). With power,
    the residue GEMMs also run back to back for about 2.5 s with the board's
    socket power and GFX clock sampled (amd-smi), beside the roofline."""
    wb = mf.ozaki_workspace_bytes(m, k, n, 0)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")

    def step():
        mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, slices, work.data_ptr(), wb, stream)
    ms = timed(torch, step, steps, warmup)
    S, fix = mf.ozaki_info(work.data_ptr(), m, k, n, slices, stream)
    n_used = mf.ozaki_moduli_used(work.data_ptr(), m, k, n, slices, stream)
    C = dC.cpu().numpy().copy()

    def step_explicit():
        mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, S, work.data_ptr(), wb, stream)
    ms_explicit = timed(torch, step_explicit, max(3, steps // 2), 1)
    N = mf.ozaki_moduli(m, k, n, S)
    res = dict(ms=ms, ms_explicit_slices=ms_explicit, slices=S, fixups=fix, moduli=n_used, moduli_explicit=N)
    if n_used:
        # the residue GEMMs of the automatic call itself (its slice and
        # moduli counts, from the control block the call left in `work`)
        step()
        ms_kern = timed(torch, lambda: mf.dev_ozaki_products(m, k, n, slices, work.data_ptr(), wb, stream),
                        max(3, steps // 2), 1)
        N = n_used
        tops = N * 2.0 * m * n * k / (ms_kern * 1e-3) / 1e12
        res["ms_kernel"] = ms_kern
        res["roofline"] = {"bound": "mfma", "achieved": round(tops, 2), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                           "frac": round(tops / INT8_PEAK_TOPS, 4),
                           "traffic": PMC_TRAFFIC.get(("k_oz2_gemm", m, S)) if m == n == k else None,
                           "note": "k_oz2_gemm: %d residue INT8 GEMMs x 2mnk int8 ops per launch (MFMA): the "
                                   "timed call's own products launch, timed alone on its stream; the rest of the call: "
                                   "exponent scan, residues with the certificate operands, certificate GEMM (FP8 "
                                   "MFMA), plans, reconstruction, fix-up" % N}
        if power:
            ms_p, pw, clk, ns = power_during(
                torch, lambda: mf.dev_ozaki_products(m, k, n, slices, work.data_ptr(), wb, stream), 2.5, 50)
            if pw is not None:
                res["roofline"]["power"] = {
                    "socket_power_w": pw, "gfx_clock_mhz": clk, "ms_per_launch": round(ms_p, 4), "samples": ns,
                    "note": "the products launch back to back for 2.5 s, amd-smi sampled alongside (medians past "
                            "the first quarter); on these random residues the board runs at its power limit and "
                            "holds the clock below 2.4 GHz (bench.py --workload products: constant residues run "
                            "the same launch at the full clock)"}
    del work
    return res, C


def bench_gemm(args, torch, mf):
    """C2 headline: float64x2 GEMM m = n = k = size, Ozaki (automatic) and plain."""
    from multifloats_amd import datagen
    m = n = k = args.size
    ri = 8
    rng = np.random.default_rng(2026)
    A = datagen.matrix(rng, m, k, args.spread)
    B = datagen.matrix(rng, k, n, args.spread)
    dA = torch.from_numpy(datagen.flat(A)).cuda()
    dB = torch.from_numpy(datagen.flat(B)).cuda()
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    flops = 2.0 * m * n * k
    es = min(args.err_sample, m)
    rows = rng.choice(m, es, replace=False)
    cols = rng.choice(n, es, replace=False)
    res = {}

    if args.form in ("both", "plain"):
        def step_plain():
            mf.dev_matmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, ri, stream, args.variant)
        ms = timed(torch, step_plain, args.steps, args.warmup)
        C = datagen.unflat(dC.cpu().numpy(), m, n)
        emax, emean = sampled_error(A, B, C, rows, cols)
        res["plain"] = dict(ms=ms, value=flops / (ms * 1e-3) / 1e9, err_max=emax, err_mean=emean,
                            roofline=gemm_roofline_plain(m, n, k, ms, args.size, args.variant))
        if m == n == k == 4096:
            # the same call back to back for about 2.5 s, amd-smi sampled alongside
            ms_p, pw, clk, ns = power_during(torch, step_plain, 2.5, 4)
            if pw is not None:
                res["plain"]["roofline"]["power"] = {
                    "socket_power_w": pw, "gfx_clock_mhz": clk, "ms_per_call": round(ms_p, 3), "samples": ns,
                    "fp64_issue_frac_at_clock": round(res["plain"]["roofline"]["fp64_issue_frac"] * 2400.0 / clk, 4),
                    "note": "the call back to back for 2.5 s with amd-smi sampled alongside (medians past the first "
                            "quarter); fp64_issue_frac_at_clock rescales the issue fraction from 2.4 GHz to the "
                            "sampled clock"}
        if not args.no_cpu:
            dt, Cc, kind = cpu_matmul(A, B, args.cpu_cols, ri)
            from tests import _native as nt
            same = bool(nt.bits_equal(Cc, C[:, :args.cpu_cols]).all())
            res["cpu"] = dict(value=round(2.0 * m * k * args.cpu_cols / dt / 1e9, 4), unit="DD-GFLOP/s", cores=1,
                              kind=kind, sample="C[:, :%d] = A(%dx%d) * B[:, :%d] on one host core, %.1f s; "
                                                "bit-identical to the GPU's plain form: %s"
                                                % (args.cpu_cols, m, k, args.cpu_cols, dt, same), **host_info())
        del C

    if args.form in ("both", "ozaki"):
        r, Cf = ozaki_run(torch, mf, dA, dB, dC, m, k, n, args.steps, args.warmup, stream, args.slices,
                          power=True)
        C = datagen.unflat(Cf, m, n)
        r["err_max"], r["err_mean"] = sampled_error(A, B, C, rows, cols)
        r["value"] = flops / (r["ms"] * 1e-3) / 1e9
        r["form"] = "crt, automatic (certified)" if args.slices <= 0 else "explicit slices"
        res["ozaki"] = r
        del C

    head = res.get("ozaki") if args.form != "plain" else res["plain"]
    ok = head["err_max"] <= 2.0 ** -100 if args.form != "plain" else True
    S = res["ozaki"]["slices"] if "ozaki" in res else 0
    metric = ("DD-GEMM DD-GFLOP/s at %d^3: Ozaki-INT8 vs plain FP64, err<=2^-100*|A||B|" % m)
    out = {
        "metric": metric,
        "value": round(head["value"], 2) if ok else None, "unit": "DD-GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8-mfma+f64" if args.form != "plain" else "f64", "data": "synthetic",
        "config": {"workload": "float64x2 GEMM m=n=k=%d, entries s*2^U(-%g,%g)*u, lo per config-0; %s"
                               % (m, args.spread, args.spread,
                                  "Ozaki INT8 automatic form (slice count chosen on the device: %d; every output "
                                  "certified to 2^-100 |A||B|) and plain reference-order (renorm 8) forms" % S
                                  if args.form == "both" else args.form),
                   "m": m, "n": n, "k": k, "slices": S, "renorm_interval": ri,
                   "error_sample": "%dx%d outputs vs exact accumulator" % (es, es)},
        "roofline": head["roofline"],
    }
    for form in ("ozaki", "plain"):
        if form in res:
            r = res[form]
            out[form] = {kk: (round(v, 4) if isinstance(v, float) and kk.startswith("ms") else
                              round(v, 2) if kk == "value" else v)
                         for kk, v in r.items() if kk != "roofline"}
            if "roofline" in r:
                out[form]["roofline"] = r["roofline"]
    if "ozaki" in res and "plain" in res:
        out["speedup_ozaki_vs_plain"] = round(res["plain"]["ms"] / res["ozaki"]["ms"], 3)
    if "cpu" in res:
        out["cpu_baseline"] = res["cpu"]
    return out


def bench_c0(args, torch, mf):
    """C0: plain GEMM 256^3, config-0 entries, renorm 8; all outputs compared
    bit for bit with the reference's own matmuldd_mm timed on one core."""
    from multifloats_amd import datagen
    from tests import _native as nt
    m = n = k = 256
    rng = np.random.default_rng(0)
    A, B = datagen.matrix(rng, m, k), datagen.matrix(rng, k, n)
    dA = torch.from_numpy(datagen.flat(A)).cuda()
    dB = torch.from_numpy(datagen.flat(B)).cuda()
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    reps = max(args.steps, 50)
    ms = timed(torch, lambda: mf.dev_matmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8, stream),
               reps, args.warmup)
    C = datagen.unflat(dC.cpu().numpy(), m, n)
    out = {"workload": "float64x2 GEMM m=n=k=256, config-0 entries (hi ~ U(-1,1)), renorm 8",
           "value": round(2.0 * m * n * k / (ms * 1e-3) / 1e9, 2), "unit": "DD-GFLOP/s", "ms": round(ms, 4),
           "steps": reps, "roofline": gemm_roofline_plain(m, n, k, ms)}
    out["roofline"]["note"] += ("; at 256^3: 16x16 tiles, one chain per lane (65,536 chains: one wave per SIMD), "
                                "256 workgroups; each lane's 256-step MAC chain is software-pipelined three steps "
                                "deep, and a lone wave per SIMD still waits on its operations' latency")
    if not args.no_cpu:
        dt, Cc, kind = cpu_matmul(A, B, n, 8, reps=3)
        out["cpu_baseline"] = dict(value=round(2.0 * m * n * k / dt / 1e9, 4), unit="DD-GFLOP/s", cores=1, kind=kind,
                                   sample="the whole 256^3 product on one host core, best of 3: %.4f s" % dt,
                                   **host_info())
        out["bit_identical_to_reference"] = bool(nt.bits_equal(Cc, C).all())
    return out


# C1: (name, op code, kind, generator, bytes per element)
C1_OPS = [("add", "ADD", "binary", "config0", 48), ("sub", "SUB", "binary", "config0", 48),
          ("mul", "MUL", "binary", "config0", 48), ("div", "DIV", "binary", "config0", 48),
          ("sqrt", "SQRT", "unary", "abs", 32), ("exp", "EXP", "unary", "exp_inputs", 32),
          ("log", "LOG", "unary", "log_inputs", 32), ("sin", "SIN", "unary", "sin_inputs", 32),
          ("erf", "ERF", "unary", "erf_inputs", 32)]
REF_LOOP_UNARY = {"sqrt": 0, "exp": 1, "log": 2, "sin": 3, "erf": 4}
REF_LOOP_BINARY = {"add": 0, "sub": 1, "mul": 2, "div": 3}


def bench_c1(args, torch, mf):
    """C1: elementwise and reductions at N = 2^26 on the C1 distributions."""
    from multifloats_amd import datagen
    from tests import _native as nt
    n = 1 << 26
    rng = np.random.default_rng(7)
    stream = torch.cuda.current_stream().cuda_stream
    _, R = ref_libs()
    res = {}
    base = datagen.config0(rng, n)
    other = datagen.config0(rng, n)
    dbase = torch.from_numpy(base).cuda()
    dother = torch.from_numpy(other).cuda()
    c = torch.empty_like(dbase)
    check = 1 << 16  # outputs compared bit for bit with the oracle
    for nm, code, kind, gen, bpe in C1_OPS:
        if gen == "config0":
            x, dx = base, dbase
        elif gen == "abs":
            x = np.where(base[:, :1] < 0, -base, base)  # |x| as a double-double
            dx = torch.from_numpy(x).cuda()
        else:
            x = getattr(datagen, gen)(rng, n)
            dx = torch.from_numpy(x).cuda()
        op = getattr(mf, "MFHIP_" + code)
        if kind == "binary":
            fn = lambda: mf.dev_binary(op, dx.data_ptr(), dother.data_ptr(), c.data_ptr(), n, stream)
        else:
            fn = lambda: mf.dev_unary(op, dx.data_ptr(), c.data_ptr(), n, stream)
        ms = timed(torch, fn, args.steps, args.warmup)
        gbs = bpe * n / ms / 1e6
        rec = {"ms": round(ms, 4), "GBs": round(gbs, 1), "Gelem_s": round(n / ms / 1e6, 2),
               "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": C1_TRAFFIC.get(nm),
                            "note": "%d algorithmic bytes per element (inputs + output)" % bpe}}
        ipe = FP64_INSTR_PER_ELEM.get(nm)
        if ipe:
            # above the FP64-issue / HBM ridge (~5 instructions per byte): the
            # FP64 pipe bounds these; the HBM figure stays beside it
            issue = ipe * n / (ms * 1e-3)
            rec["hbm_frac"] = rec["roofline"]["frac"]
            rec["roofline"] = {"bound": "fp64", "achieved": round(issue / 1e12, 3),
                               "peak": round(FP64_ISSUE_PEAK / 1e12, 1), "unit": "T FP64 instructions/s",
                               "frac": round(issue / FP64_ISSUE_PEAK, 4), "traffic": C1_TRAFFIC.get(nm),
                               "note": "%.1f FP64 ADD/MUL/FMA lane-instructions per element (PMC, this input "
                                       "distribution) x elements/s, against 256 CUs x 64 lanes x 2.4 GHz" % ipe}
            rec["fp64_issue_frac"] = rec["roofline"]["frac"]
        got = c[:check].cpu().numpy()
        if nm != "erf":  # bit-identical to the reference (erf: own design, accuracy in the tests)
            want = nt.or_binary(nm, x[:check], other[:check]) if kind == "binary" else nt.or_unary(nm, x[:check])
            rec["bit_identical_sample"] = bool(nt.bits_equal(got, want).all())
        if not args.no_cpu and R is not None:
            ns = (1 << 22) if kind == "binary" or nm == "sqrt" else (1 << 17)
            xs = np.ascontiguousarray(x[:ns])
            ys = np.ascontiguousarray(other[:ns])
            cs = np.empty_like(xs)
            t0 = time.perf_counter()
            if kind == "binary":
                R.refloop_binary(REF_LOOP_BINARY[nm], xs.ctypes.data, ys.ctypes.data, cs.ctypes.data, ns)
            else:
                R.refloop_unary(REF_LOOP_UNARY[nm], xs.ctypes.data, cs.ctypes.data, ns)
            dt = time.perf_counter() - t0
            rec["cpu_baseline"] = {"value": round(ns / dt / 1e9, 5), "unit": "G elements/s", "cores": 1,
                                   "kind": "reference", "ns_per_elem": round(dt / ns * 1e9, 2),
                                   "sample": "%d elements of the same inputs through the reference's %sdd on one "
                                             "host core, %.2f s" % (ns, nm, dt)}
            rec["speedup_vs_cpu_core"] = round((n / ms / 1e6) / (ns / dt / 1e9), 1)
        res[nm] = rec
        if gen not in ("config0",):
            del dx
    del dbase, c
    # reductions: dot on 10% near-cancelling pairs, sum on near-cancelling
    # neighbours, norm2 on config-0 entries
    a_np, b_np = datagen.cancel_pairs(rng, n)
    s_np = datagen.sum_inputs(rng, n)
    work = torch.empty(mf.reduce_workspace_bytes(n) // 8 + 2, dtype=torch.float64, device="cuda")
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    O = nt.oracle()
    for nm, kind, bpe, xs in (("dot", mf.MFHIP_DOT, 32, (a_np, b_np)), ("sum", mf.MFHIP_SUM, 16, (s_np, s_np)),
                              ("norm2", mf.MFHIP_NORM2, 16, (other, other))):
        da = torch.from_numpy(xs[0]).cuda()
        db = torch.from_numpy(xs[1]).cuda() if nm == "dot" else da
        fn = lambda: mf.dev_reduce(kind, da.data_ptr(), db.data_ptr(), n, out.data_ptr(), work.data_ptr(), 8, stream)
        ms = timed(torch, fn, args.steps, args.warmup)
        gbs = bpe * n / ms / 1e6
        val = out.cpu().numpy()
        rec = {"ms": round(ms, 4), "GBs": round(gbs, 1),
               "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": C1_TRAFFIC.get(nm),
                            "note": "%d algorithmic bytes per element (first pass; the fold passes read "
                                    "one partial per 2,048 or 4,096 elements)" % bpe}}
        if nm == "dot":
            e, mg = nt.exact_dot_err(xs[0], xs[1], val)
            rec["err_over_n_sum_abs_terms"] = e / mg / n
        elif nm == "sum":
            e, mg = nt.exact_sum_err(xs[0], val)
            rec["err_over_n_sum_abs_terms"] = e / mg / n
        if not args.no_cpu:
            ns = 1 << 22
            p0 = np.ascontiguousarray(xs[0][:ns])
            p1 = np.ascontiguousarray(xs[1][:ns])
            t0 = time.perf_counter()
            if nm == "dot":
                O.or_dot(nt.ptr(p0), nt.ptr(p1), ns, 8)
            elif nm == "sum":
                O.or_sum(nt.ptr(p0), ns)
            else:
                O.or_norm2(nt.ptr(p0), ns)
            dt = time.perf_counter() - t0
            rec["cpu_baseline"] = {"value": round(bpe * ns / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                                   "sample": "%d elements through the oracle's restatement of the Fortran %s loop "
                                             "(the reference has no C entry point for it) on one host core, %.3f s"
                                             % (ns, nm, dt)}
        res[nm] = rec
        del da, db
    hbm_ops = [k_ for k_ in res if res[k_]["roofline"]["bound"] == "hbm"]
    worst = min(hbm_ops, key=lambda k_: res[k_]["GBs"])
    out = {"workload": "float64x2 arrays N=2^26 per op: + - * / on config-0 pairs, sqrt on |x|, exp on U[-700,700], "
                       "log on 10^U(-300,300), sin on U[-1e4,1e4] with 1%% near k pi/2, erf on U[-6,6]; dot on 10%% "
                       "near-cancelling pairs, sum on near-cancelling neighbours, norm2 on config-0",
           "value": res[worst]["GBs"], "unit": "GB/s", "value_is": "the slowest HBM-bound op (%s)" % worst,
           "roofline": res[worst]["roofline"], "ops": res}
    if "cpu_baseline" in res[worst]:
        out["cpu_baseline"] = dict(res[worst]["cpu_baseline"], **host_info())
    return out


def bench_c2_complex(args, torch, mf):
    """complex64x2 GEMM m = n = k = 2048, spread 2^+-20, renorm 8; a sampled
    block compared bit for bit with the oracle's four-GEMM composition."""
    from multifloats_amd import datagen
    from tests import _native as nt
    m = n = k = 2048
    rng = np.random.default_rng(3)
    re_a, im_a = datagen.matrix(rng, m, k, 20), datagen.matrix(rng, m, k, 20)
    re_b, im_b = datagen.matrix(rng, k, n, 20), datagen.matrix(rng, k, n, 20)

    def pack(re, im):
        return np.ascontiguousarray(np.concatenate([re, im], axis=2).transpose(1, 0, 2)).reshape(-1, 4)
    dA, dB = torch.from_numpy(pack(re_a, im_a)).cuda(), torch.from_numpy(pack(re_b, im_b)).cuda()
    dC = torch.empty((m * n, 4), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    ms = timed(torch, lambda: mf.dev_cmatmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8, stream),
               max(3, args.steps // 2), args.warmup)
    C = dC.cpu().numpy().reshape(n, m, 4).transpose(1, 0, 2)
    rows = rng.choice(m, 8, replace=False)
    cols = rng.choice(n, 8, replace=False)
    sa = np.concatenate([re_a[rows], im_a[rows]], axis=2)
    sb = np.concatenate([re_b[:, cols], im_b[:, cols]], axis=2)
    O = nt.oracle()
    Ac = np.ascontiguousarray(sa.transpose(1, 0, 2))
    Bc = np.ascontiguousarray(sb.transpose(1, 0, 2))
    Wc = np.zeros((8, 8, 4))
    O.or_cmatmul_mm(nt.ptr(Ac), nt.ptr(Bc), nt.ptr(Wc), 8, k, 8, 8)
    want = Wc.transpose(1, 0, 2)
    got = C[np.ix_(rows, cols)]
    achieved = 4 * FP64_FLOPS_PER_DDMAC * m * n * k / (ms * 1e-3) / 1e12
    return {"workload": "complex64x2 GEMM m=n=k=2048, spread 2^+-20, renorm 8 (fused: four reference-order chains "
                        "per output)", "value": round(8.0 * m * n * k / (ms * 1e-3) / 1e9, 2), "unit": "DD-GFLOP/s",
            "ms": round(ms, 4),
            "roofline": {"bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None},
            "bit_identical_sample_8x8": bool(nt.bits_equal(got, want).all())}


def bench_c3(args, torch, mf):
    """C3: Ozaki 8192^3 (automatic form) at exponent spreads 2^+-10 and 2^+-30,
    with the plain reference-order GEMM at 8192^3 (spread 2^+-10) beside it;
    error max / mean on a 64 x 64 block of sampled outputs vs the exact
    accumulator."""
    from multifloats_amd import datagen
    m = n = k = 8192
    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": "float64x2 GEMM m=n=k=8192, entries s*2^U(-e,e)*u (e = 10, 30), lo per config-0; Ozaki INT8 "
                       "automatic form; plain reference-order form beside it (e = 10)"}
    flops = 2.0 * m * n * k
    per = {}
    for spread in (10, 30):
        rng = np.random.default_rng(8192 + spread)
        A = datagen.matrix(rng, m, k, spread)
        B = datagen.matrix(rng, k, n, spread)
        dA = torch.from_numpy(datagen.flat(A)).cuda()
        dB = torch.from_numpy(datagen.flat(B)).cuda()
        dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
        r, Cf = ozaki_run(torch, mf, dA, dB, dC, m, k, n, args.steps, args.warmup, stream)
        rows = rng.choice(m, 64, replace=False)
        cols = rng.choice(n, 64, replace=False)
        C = datagen.unflat(Cf, m, n)
        r["err_max"], r["err_mean"] = sampled_error(A, B, C, rows, cols)
        r["value"] = round(flops / (r["ms"] * 1e-3) / 1e9, 2)
        r = {kk: (round(v, 4) if isinstance(v, float) and kk.startswith("ms") else v) for kk, v in r.items()}
        per["spread_%d" % spread] = r
        if spread == 10:
            steps_p = min(args.steps, 3)
            ms_p = timed(torch, lambda: mf.dev_matmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8, stream),
                         steps_p, 1)
            Cp = datagen.unflat(dC.cpu().numpy(), m, n)
            emax, emean = sampled_error(A, B, Cp, rows, cols)
            out["plain"] = {"ms": round(ms_p, 3), "steps": steps_p, "value": round(flops / (ms_p * 1e-3) / 1e9, 2),
                            "err_max": emax, "err_mean": emean, "roofline": gemm_roofline_plain(m, n, k, ms_p)}
            if not args.no_cpu:
                cols_cpu = 4
                dt, Cc, kind = cpu_matmul(A, B, cols_cpu, 8)
                from tests import _native as nt
                same = bool(nt.bits_equal(Cc, Cp[:, :cols_cpu]).all())
                out["cpu_baseline"] = dict(value=round(2.0 * m * k * cols_cpu / dt / 1e9, 4), unit="DD-GFLOP/s",
                                           cores=1, kind=kind,
                                           sample="C[:, :%d] = A(8192x8192) * B[:, :%d] (spread 2^+-10) on one host "
                                                  "core, %.1f s; the whole product at that rate: %.0f s; "
                                                  "bit-identical to the GPU's plain form: %s"
                                                  % (cols_cpu, cols_cpu, dt, dt * n / cols_cpu, same), **host_info())
            del Cp
        del A, B, C, dA, dB, dC
        torch.cuda.empty_cache()
    out["ozaki"] = per
    head = per["spread_10"]
    out["value"] = head["value"]
    out["unit"] = "DD-GFLOP/s"
    out["value_is"] = "Ozaki at spread 2^+-10"
    out["roofline"] = head.get("roofline")
    out["speedup_ozaki_vs_plain"] = {kk: round(out["plain"]["ms"] / v["ms"], 2) for kk, v in per.items()}
    out["within_2^-100"] = all(v["err_max"] <= 2.0 ** -100 for v in per.values())
    return out


def bench_c4(args, torch, mf):
    """C4: the batching form, with a CPU baseline of the same problems."""
    from multifloats_amd import datagen
    from tests import _native as nt
    out, extra = bench_batched(args, torch, mf, keep=True)
    if not args.no_cpu:
        sizes, dlen, gb, db, offs, doff = extra
        rng = np.random.default_rng(44)
        O = nt.oracle()
        t_g = t_d = 0.0
        macs = elems = 0
        for i in rng.choice(len(sizes), 150, replace=False):
            m, k, n = (int(v) for v in sizes[i])
            o = offs[i]
            A = np.ascontiguousarray(gb[o:o + m * k].cpu().numpy())
            B = np.ascontiguousarray(gb[o + m * k:o + m * k + k * n].cpu().numpy())
            C = np.zeros((m * n, 2))
            t0 = time.perf_counter()
            O.or_matmul_mm(nt.ptr(A), nt.ptr(B), nt.ptr(C), m, k, n, 8)
            t_g += time.perf_counter() - t0
            macs += m * k * n
        for i in rng.choice(len(dlen), 600, replace=False):
            L, o = int(dlen[i]), doff[i]
            a = np.ascontiguousarray(db[o:o + L].cpu().numpy())
            b = np.ascontiguousarray(db[o + L:o + 2 * L].cpu().numpy())
            t0 = time.perf_counter()
            O.or_dot(nt.ptr(a), nt.ptr(b), L, 8)
            t_d += time.perf_counter() - t0
            elems += L
        rate = (2.0 * macs + 2.0 * elems) / (t_g + t_d) / 1e9
        out["cpu_baseline"] = dict(value=round(rate, 4), unit="DD-GFLOP/s", cores=1, kind="port",
                                   sample="150 of the GEMMs (%.3g MACs, %.2f s) and 600 of the dots (%d elements, "
                                          "%.2f s) through the oracle's restatements of matmuldd_mm and the Fortran "
                                          "dot_product, one problem after another on one host core"
                                          % (macs, t_g, elems, t_d), **host_info())
    return out


def bench_all(args, torch, mf):
    t0 = time.perf_counter()
    head = bench_gemm(args, torch, mf)
    configs = {}
    timing = {}

    def run(name, fn):
        t = time.perf_counter()
        print("bench: %s ..." % name, file=sys.stderr, flush=True)
        configs[name] = fn(args, torch, mf)
        timing[name] = round(time.perf_counter() - t, 1)
        torch.cuda.empty_cache()
    c2 = {"workload": head["config"]["workload"], "value": head["value"], "unit": "DD-GFLOP/s",
          "roofline": head["roofline"], "ozaki": head.get("ozaki"), "plain": head.get("plain"),
          "speedup_ozaki_vs_plain": head.get("speedup_ozaki_vs_plain")}
    if "cpu_baseline" in head:
        c2["cpu_baseline"] = head["cpu_baseline"]
    run("C0", bench_c0)
    run("C1", bench_c1)
    configs["C2"] = c2
    run("C2_complex", bench_c2_complex)
    configs["C2"]["complex"] = configs.pop("C2_complex")
    if not args.quick:
        run("C3", bench_c3)
        run("C4", bench_c4)
    head["configs"] = {k_: configs[k_] for k_ in sorted(configs)}
    head["bench_seconds"] = dict(timing, total=round(time.perf_counter() - t0, 1))
    return head


def bench_blas(args, torch, mf):
    """BLAS-style C = alpha op(A) op(B) + beta C at m = n = k = size for the four
    transpose pairs (general alpha and beta, renorm interval 8)."""
    from multifloats_amd import datagen
    m = n = k = args.size
    rng = np.random.default_rng(5)
    dev = {}
    for nm in ("A", "B", "C"):
        dev[nm] = torch.from_numpy(np.ascontiguousarray(
            datagen.matrix(rng, m, m, args.spread).transpose(1, 0, 2)).reshape(-1, 2)).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    alpha, beta = (1.5, 2.0 ** -60), (-0.75, 2.0 ** -58)
    per = {}
    for ta, tb in ("NN", "TN", "NT", "TT"):
        def step():
            mf.dev_gemm(ta, tb, m, n, k, alpha, dev["A"].data_ptr(), m, dev["B"].data_ptr(), k, beta,
                        dev["C"].data_ptr(), m, 8, stream)
        ms = timed(torch, step, args.steps, args.warmup)
        per[ta + tb] = {"ms": round(ms, 4), "dd_gflops": round(2.0 * m * n * k / (ms * 1e-3) / 1e9, 2)}
    worst = max(per.values(), key=lambda r: r["ms"])
    achieved = FP64_FLOPS_PER_DDMAC * m * n * k / (worst["ms"] * 1e-3) / 1e12
    return {"metric": "BLAS-style DD-GEMM DD-GFLOP/s (slowest transpose pair)", "value": worst["dd_gflops"],
            "unit": "DD-GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": worst["ms"], "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "float64x2 C = alpha op(A) op(B) + beta C, m=n=k=%d, NN/TN/NT/TT, "
                                   "renorm 8" % m, "per_pair": per},
            "roofline": {"bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None},
            "vs_baseline": None}


FP64_FLOPS_PER_TRSM_UPDATE = 29  # x - c*y: DD mul 9 + DD sub 20


def bench_trsm(args, torch, mf):
    """Triangular solves B <- op(A)^-1 alpha B / alpha B op(A)^-1, m = n = size,
    all eight shim variants (non-unit diagonal). The six blocked variants run at
    the full size; the two whose sums run against the pivot order (L/L/T,
    R/L/N) are one dependent chain per vector and run at size / 4."""
    from multifloats_amd import datagen
    size = args.size
    rng = np.random.default_rng(9)
    stream = torch.cuda.current_stream().cuda_stream
    per, blocked = {}, []
    for side, uplo, ta in [(a, b, c) for a in "LR" for b in "UL" for c in "NT"]:
        chain = uplo == "L" and ((side == "L") == (ta == "T"))
        s = size // 4 if chain else size
        A = datagen.matrix(rng, s, s)
        A[np.arange(s), np.arange(s)] += np.array([4.0, 0.0])  # well-conditioned diagonal
        dA = torch.from_numpy(np.ascontiguousarray(A.transpose(1, 0, 2))).cuda()
        dB = torch.from_numpy(np.ascontiguousarray(datagen.matrix(rng, s, s).transpose(1, 0, 2))).cuda()
        wb = mf.trsm_workspace_bytes(side, s, s)
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        # every step solves the same system again (B is restored from a copy first)
        dB0 = dB.clone()

        def step():
            dB.copy_(dB0)
            mf.dev_trsm(side, uplo, ta, "N", s, s, 1.5, dA.data_ptr(), s, dB.data_ptr(), s, work.data_ptr(), wb,
                        stream)
        copy_ms = timed(torch, lambda: dB.copy_(dB0), args.steps, args.warmup)
        ms = timed(torch, step, args.steps, args.warmup) - copy_ms
        updates = s * (s - 1) / 2.0 * s
        rec = {"size": s, "ms": round(ms, 4), "dd_gflops": round(2.0 * updates / (ms * 1e-3) / 1e9, 2)}
        per[side + uplo + ta] = rec
        if not chain:
            blocked.append(rec)
    worst = min(blocked, key=lambda r: r["dd_gflops"])
    achieved = FP64_FLOPS_PER_TRSM_UPDATE * worst["dd_gflops"] / 2.0 / 1e3
    return {"metric": "DD-TRSM DD-GFLOP/s (slowest blocked variant)", "value": worst["dd_gflops"],
            "unit": "DD-GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": worst["ms"], "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "float64x2 triangular solve m=n=%d (chain variants %d), alpha 1.5, non-unit, "
                                   "bit-identical to the xTRSM shim" % (size, size // 4), "per_variant": per},
            "roofline": {"bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None},
            "vs_baseline": None}


def bench_batched(args, torch, mf, keep=False):
    """10,000 float64x2 GEMMs (m, n, k log-uniform in [8, 512]) and 50,000
    dot products (length log-uniform in [16, 65536]), config-0 entries, in one
    persistent grouped launch replayed from a HIP graph; per-call launches of
    the same problems as the comparison."""
    from multifloats_amd import datagen, autotune
    from tests import _native as nt  # oracle: accuracy of sampled problems
    rng = np.random.default_rng(4)
    sizes, dlen = datagen.batched_sizes(rng)
    table = autotune.load_table()
    if table is None or args.retune:
        table = autotune.tune(torch, mf)
    variants = autotune.plan_variants(sizes, table)
    tot = int(sum(m * k + k * n + m * n for m, k, n in sizes))
    gbuf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, gbuf, 101)
    dtot = int(dlen.sum())
    dbuf = torch.empty(4 * dtot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, dbuf, 202)
    dout = torch.zeros((len(dlen), 2), dtype=torch.float64, device="cuda")
    base, dbase = gbuf.data_ptr(), dbuf.data_ptr()
    gemms, offs, off = [], [], 0
    for m, k, n in sizes:
        m, k, n = int(m), int(k), int(n)
        offs.append(off)
        gemms.append((base + 16 * off, base + 16 * (off + m * k), base + 16 * (off + m * k + k * n), m, k, n, 8))
        off += m * k + k * n + m * n
    dots, doff, off = [], [], 0
    for i, L in enumerate(dlen):
        L = int(L)
        doff.append(off)
        dots.append((dbase + 16 * off, dbase + 16 * (off + L), dout.data_ptr() + 16 * i, L, 8))
        off += 2 * L
    plan = mf.GroupedPlan(gemms, dots, variants)
    stream = torch.cuda.current_stream().cuda_stream
    ms = timed(torch, lambda: plan.run_graph(stream), args.steps, args.warmup)
    ms_eager = timed(torch, lambda: plan.run(stream), max(1, args.steps // 2), 1)
    macs = float(sum(int(m) * int(k) * int(n) for m, k, n in sizes))
    dd_flops = 2.0 * macs + 2.0 * float(dtot)

    # per-call launches of the same problems (one stream, the library's own kernels)
    wbytes = mf.reduce_workspace_bytes(int(dlen.max()))
    work = torch.empty(wbytes // 8 + 2, dtype=torch.float64, device="cuda")
    pout = torch.zeros((len(dlen), 2), dtype=torch.float64, device="cuda")
    def per_call():
        for g in gemms:
            mf.dev_matmul(g[0], g[1], g[2], g[3], g[4], g[5], 8, stream)
        for i, d in enumerate(dots):
            mf.dev_reduce(mf.MFHIP_DOT, d[0], d[1], d[3], pout.data_ptr() + 16 * i, work.data_ptr(), 8, stream)
    ms_calls = timed(torch, per_call, 1, 1)

    # accuracy: sampled GEMMs bit for bit against the oracle, sampled dots
    # against the exact accumulator
    gb = gbuf.view(-1, 2)
    bits_ok = True
    gerr = 0.0
    for i in rng.choice(len(sizes), 12, replace=False):
        m, k, n = (int(v) for v in sizes[i])
        o = offs[i]
        A = datagen.unflat(gb[o:o + m * k].cpu().numpy(), m, k)
        B = datagen.unflat(gb[o + m * k:o + m * k + k * n].cpu().numpy(), k, n)
        C = datagen.unflat(gb[o + m * k + k * n:o + m * k + k * n + m * n].cpu().numpy(), m, n)
        bits_ok &= bool(nt.bits_equal(C, nt.or_matmul(A, B, 8)).all())
        rr, cc = np.meshgrid(np.arange(min(m, 8)), np.arange(min(n, 8)), indexing="ij")
        e, mg = nt.gemm_err(A, B, C, rr.ravel(), cc.ravel())
        gerr = max(gerr, float((e / mg).max()))
    db = dbuf.view(-1, 2)
    dres = dout.cpu().numpy()
    derr = 0.0
    for i in rng.choice(len(dlen), 40, replace=False):
        L, o = int(dlen[i]), doff[i]
        a, b = db[o:o + L].cpu().numpy(), db[o + L:o + 2 * L].cpu().numpy()
        e, mg = nt.exact_dot_err(a, b, dres[i])
        derr = max(derr, e / mg / L)
    plan.close()
    out = {"metric": "batched DD GEMMs + dots, aggregate DD-GFLOP/s", "value": round(dd_flops / (ms * 1e-3) / 1e9, 2),
            "unit": "DD-GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 4), "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "10000 float64x2 GEMMs m,n,k log-U[8,512] + 50000 dots len log-U[16,65536], "
                                   "config-0 entries, renorm 8; one persistent grouped launch in a HIP graph",
                       "gemm_macs": macs, "dot_elements": dtot, "work_items": plan.items,
                       "tile_table": table},
            "graph_ms": round(ms, 4), "eager_grouped_ms": round(ms_eager, 4), "per_call_ms": round(ms_calls, 3),
            "speedup_vs_per_call": round(ms_calls / ms, 2),
            "gemm_dd_gflops": round(2.0 * macs / (ms * 1e-3) / 1e9, 2),
            "accuracy": {"sampled_gemms_bit_identical_to_reference_order": bits_ok,
                         "gemm_max_err_over_absAB": gerr, "dot_max_err_over_n_sum_abs": derr}}
    gemm_ops = 2.0 * macs / (ms * 1e-3) / 1e12 * FP64_FLOPS_PER_DDMAC / 2.0
    out["roofline"] = {"bound": "fp64", "achieved": round(gemm_ops, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(gemm_ops / FP64_PEAK_TFLOPS, 4), "traffic": None,
                       "note": "the grouped kernel's GEMM work at %.2f FP64 flops per DD-MAC over the whole graph "
                               "replay (the dots' %.2f GB of HBM reads share the launch)"
                               % (FP64_FLOPS_PER_DDMAC, 32.0 * dtot / 1e9)}
    if keep:
        return out, (sizes, dlen, gb, db, offs, doff)
    return out




def bench_elementwise(args, torch, mf):
    """Every array function at N = 2^26 (C1 covers the hot-path ones)."""
    from multifloats_amd import datagen
    n = 1 << 26
    rng = np.random.default_rng(7)
    stream = torch.cuda.current_stream().cuda_stream
    res = {}
    c = torch.empty((n, 2), dtype=torch.float64, device="cuda")
    gens = {"exp": datagen.exp_inputs, "log": datagen.log_inputs, "sin": datagen.sin_inputs, "erf": datagen.erf_inputs,
            "cos": datagen.sin_inputs, "tan": datagen.sin_inputs, "exp2": datagen.exp_inputs,
            "expm1": datagen.config0, "log2": datagen.log_inputs, "log10": datagen.log_inputs,
            "log1p": datagen.config0, "sinh": datagen.erf_inputs, "cosh": datagen.erf_inputs,
            "tanh": datagen.erf_inputs, "erfc": datagen.erf_inputs, "erfcx": datagen.erf_inputs,
            "asinh": datagen.erf_inputs, "acosh": datagen.acosh_inputs, "atanh": datagen.atanh_inputs,
            "exp10": datagen.exp10_inputs, "exp2m1": datagen.exp_inputs, "exp10m1": datagen.exp10_inputs,
            "log2p1": datagen.config0, "log10p1": datagen.config0, "sinpi": datagen.erf_inputs,
            "cospi": datagen.erf_inputs, "tanpi": datagen.erf_inputs, "atan": datagen.erf_inputs,
            "asin": datagen.atanh_inputs, "acos": datagen.atanh_inputs}
    for nm, gen in gens.items():
        x = torch.from_numpy(gen(rng, n)).cuda()
        op = getattr(mf, "MFHIP_" + nm.upper())
        ms = timed(torch, lambda: mf.dev_unary(op, x.data_ptr(), c.data_ptr(), n, stream), args.steps, args.warmup)
        res[nm] = dict(ms=round(ms, 4), GBs=round(32.0 * n / ms / 1e6, 1), Gelem_s=round(n / ms / 1e6, 2))
        del x
    worst = min(res, key=lambda k_: res[k_]["Gelem_s"])
    return {"metric": "elementwise DD array functions, N=2^26, G elements/s (slowest function)",
            "value": res[worst]["Gelem_s"], "unit": "G elements/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "every one-argument array function on float64x2, N=2^26", "slowest": worst},
            "ops": res}


def bench_profile(args, torch, mf):
    """One launch of each hot kernel at its full size, for rocprofv3 runs
    (--kernel-trace --stats, or --pmc counter passes): a 1 GiB device copy
    (the calibration of the FETCH_SIZE / WRITE_SIZE counters), the Ozaki
    automatic call and the plain GEMM at 4096^3 (spread 2^+-20), + - * / sqrt
    exp log sin erf on the C1 inputs, dot / sum / norm2 at N = 2^26, and the
    C4 mix (10,000 GEMMs + 50,000 dots) through the grouped plan."""
    from multifloats_amd import datagen
    s = torch.cuda.current_stream().cuda_stream
    x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    y.copy_(x)  # 1 GiB read + 1 GiB written
    del x, y
    m = n = k = 4096
    rng = np.random.default_rng(2026)
    dA = torch.from_numpy(datagen.flat(datagen.matrix(rng, m, k, 20))).cuda()
    dB = torch.from_numpy(datagen.flat(datagen.matrix(rng, k, n, 20))).cuda()
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    wb = mf.ozaki_workspace_bytes(m, k, n, 0)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, s)
    mf.dev_matmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8, s)
    torch.cuda.synchronize()
    del dA, dB, dC, work
    n1 = 1 << 26
    rng = np.random.default_rng(7)
    a = torch.from_numpy(datagen.config0(rng, n1)).cuda()
    b = torch.from_numpy(datagen.config0(rng, n1)).cuda()
    c = torch.empty_like(a)
    for code in ("ADD", "SUB", "MUL", "DIV"):
        mf.dev_binary(getattr(mf, "MFHIP_" + code), a.data_ptr(), b.data_ptr(), c.data_ptr(), n1, s)
    a.abs_()
    mf.dev_unary(mf.MFHIP_SQRT, a.data_ptr(), c.data_ptr(), n1, s)
    for nm in ("exp", "log", "sin", "erf"):
        x = torch.from_numpy(getattr(datagen, nm + "_inputs")(rng, n1)).cuda()
        mf.dev_unary(getattr(mf, "MFHIP_" + nm.upper()), x.data_ptr(), c.data_ptr(), n1, s)
        torch.cuda.synchronize()
        del x
    work = torch.empty(mf.reduce_workspace_bytes(n1) // 8 + 2, dtype=torch.float64, device="cuda")
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    for kind in (mf.MFHIP_DOT, mf.MFHIP_SUM, mf.MFHIP_NORM2):
        mf.dev_reduce(kind, a.data_ptr(), b.data_ptr(), n1, out.data_ptr(), work.data_ptr(), 8, s)
    torch.cuda.synchronize()
    del a, b, c, work
    # the C4 mix: one eager run of the grouped plan (one launch per tile kind, then the dots)
    from multifloats_amd import autotune
    sizes, dlen = datagen.batched_sizes(np.random.default_rng(4))
    tot = int(sum(m * k + k * n + m * n for m, k, n in sizes))
    gbuf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, gbuf, 101)
    dtot = int(dlen.sum())
    dbuf = torch.empty(4 * dtot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, dbuf, 202)
    dout = torch.zeros((len(dlen), 2), dtype=torch.float64, device="cuda")
    gemms, off = [], 0
    for m, k, n in sizes:
        m, k, n = int(m), int(k), int(n)
        g0 = gbuf.data_ptr() + 16 * off
        gemms.append((g0, g0 + 16 * m * k, g0 + 16 * (m * k + k * n), m, k, n, 8))
        off += m * k + k * n + m * n
    dots, off = [], 0
    for i, L in enumerate(dlen):
        L = int(L)
        d0 = dbuf.data_ptr() + 16 * off
        dots.append((d0, d0 + 16 * L, dout.data_ptr() + 16 * i, L, 8))
        off += 2 * L
    plan = mf.GroupedPlan(gemms, dots, autotune.plan_variants(sizes))
    plan.run(s)
    torch.cuda.synchronize()
    plan.close()
    return {"metric": "profile run (one launch of each hot kernel)", "value": None, "unit": None}


def bench_int8ref(args, torch, mf):
    """The vendor's INT8 GEMM (hipBLASLt through torch._int_mm: int8 x int8
    -> int32, B column-major, its faster layout here) at m = n = k = 4096 and
    8192 on uniformly random bytes (what the Ozaki residues are), on bytes in
    [-4, 3] and on constant bytes: the rate the matrix cores reach on each kind
    of data, the yardstick for k_oz2_gemm."""
    out = {}
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    for size in (4096, 8192):
        row = {}
        for kind in ("random", "small", "constant"):
            lo, hi = {"random": (-128, 128), "small": (-4, 4), "constant": (3, 4)}[kind]
            a = torch.randint(lo, hi, (size, size), dtype=torch.int8, device="cuda", generator=g)
            b = torch.randint(lo, hi, (size, size), dtype=torch.int8, device="cuda", generator=g).t().contiguous().t()
            ms = timed(torch, lambda: torch._int_mm(a, b), args.steps, args.warmup)
            row[kind] = {"ms": round(ms, 4), "TOPs": round(2.0 * size ** 3 / ms / 1e9, 1)}
            del a, b
        out[str(size)] = row
    return {"metric": "INT8 GEMM rate of the vendor library (hipBLASLt via torch._int_mm) by operand data",
            "value": out["4096"]["random"]["TOPs"], "unit": "TOP/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "dtype": "int8", "data": "synthetic",
            "config": {"workload": "int8 x int8 -> int32 GEMM, m = n = k in {4096, 8192}; random, [-4, 3], "
                                   "constant bytes"}, "rates": out, "peak": INT8_PEAK_TOPS}


def _power_sampler(samples, stop):
    """Socket power (W) and GFX clock (MHz) from amd-smi, every ~0.3 s until
    `stop` is set; nothing is recorded where amd-smi is absent."""
    import re
    import subprocess
    while not stop.is_set():
        try:
            out = subprocess.run(["amd-smi", "metric", "-p", "-c"], capture_output=True, text=True,
                                 timeout=5).stdout
        except (OSError, subprocess.SubprocessError):
            return
        p = re.search(r"SOCKET_POWER:\s*(\d+)", out)
        c = re.search(r"GFX_0:\s*\n\s*CLK:\s*(\d+)\s*MHz", out)
        if p and c:
            samples.append((int(p.group(1)), int(c.group(1))))
        stop.wait(0.3)


def power_during(torch, fn, seconds=3.0, batch=100):
    """Run fn back to back for about `seconds` in batches under the amd-smi
    sampler; returns (median ms per call past the first quarter of the
    batches, median socket power W, median GFX clock MHz, samples), the last
    three None where amd-smi is absent."""
    import threading
    samples, stop = [], threading.Event()
    th = threading.Thread(target=_power_sampler, args=(samples, stop), daemon=True)
    th.start()
    batches = []
    t_end = time.time() + seconds
    while time.time() < t_end:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(batch):
            fn()
        e1.record()
        e1.synchronize()
        batches.append(e0.elapsed_time(e1) / batch)
    stop.set()
    th.join()
    tail = batches[len(batches) // 4:] or batches
    ms = sorted(tail)[len(tail) // 2]
    if not samples:
        return ms, None, None, 0
    body = samples[len(samples) // 4:] or samples  # past the ramp-up
    return (ms, sorted(s[0] for s in body)[len(body) // 2], sorted(s[1] for s in body)[len(body) // 2],
            len(samples))


def bench_products(args, torch, mf):
    """The products stage of an automatic Ozaki call (k_oz2_gemm, the residue
    INT8 GEMMs, timed alone) by operand data, with the part's socket power and
    GFX clock sampled while it runs: the bench's iid operands (spread 2^+-20:
    uniformly random residues) against A = B = all ones (every residue plane
    constant). Same launch, same cycle count; the clock the part holds under
    each load sets the rate."""
    from multifloats_amd import datagen
    m = n = k = args.size
    rng = np.random.default_rng(2026)
    stream = torch.cuda.current_stream().cuda_stream
    wb = mf.ozaki_workspace_bytes(m, k, n, 0)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    out = {}
    for kind in ("random", "constant"):
        if kind == "random":
            dA = torch.from_numpy(datagen.flat(datagen.matrix(rng, m, k, args.spread))).cuda()
            dB = torch.from_numpy(datagen.flat(datagen.matrix(rng, k, n, args.spread))).cuda()
        else:
            dA = torch.zeros((m * k, 2), dtype=torch.float64, device="cuda")
            dA[:, 0] = 1.0
            dB = torch.zeros((k * n, 2), dtype=torch.float64, device="cuda")
            dB[:, 0] = 1.0
        mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, stream)
        N = mf.ozaki_moduli_used(work.data_ptr(), m, k, n, 0, stream)

        def prod():
            mf.dev_ozaki_products(m, k, n, 0, work.data_ptr(), wb, stream)
        for _ in range(args.warmup):
            prod()
        # about 3 s of back-to-back launches in batches of 100 under the
        # sampler; the rate is the median batch past the first quarter
        t0 = time.time()
        ms, pw, clk, ns = power_during(torch, prod, 3.0, 100)
        row = {"moduli": N, "ms": round(ms, 4), "TOPs": round(N * 2.0 * m * n * k / ms / 1e9, 1),
               "launches": int(round((time.time() - t0) * 1e3 / ms))}
        if pw is not None:
            row.update(socket_power_w=pw, gfx_clock_mhz=clk, mcycles_per_launch=round(ms * 1e-3 * clk * 1e6 / 1e6, 3),
                       samples=ns)
        out[kind] = row
        del dA, dB
    del work
    return {"metric": "Ozaki products stage (k_oz2_gemm) rate by operand data, with socket power and GFX clock",
            "value": out["random"]["TOPs"], "unit": "TOP/s", "n_gpus": 1, "steps": out["random"]["launches"],
            "warmup": args.warmup, "higher_is_better": True, "dtype": "int8", "data": "synthetic",
            "config": {"workload": "residue INT8 GEMMs of an automatic %d^3 call: iid spread 2^+-%g operands "
                                   "(random residues) and all-ones operands (constant residue planes)"
                                   % (m, args.spread)},
            "cases": out, "peak": INT8_PEAK_TOPS}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the DD kernels run on a single MI355X here; "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    import torch
    import multifloats_amd as mf
    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    fn = {"all": bench_all, "gemm": bench_gemm, "cgemm": lambda a, t, m: bench_c2_complex(a, t, m),
          "blas": bench_blas, "trsm": bench_trsm, "batched": bench_batched, "elementwise": bench_elementwise,
          "reduce": lambda a, t, m: bench_c1(a, t, m), "profile": bench_profile, "int8ref": bench_int8ref,
          "products": bench_products}[args.workload]
    print(json.dumps(fn(args, torch, mf)), flush=True)


if __name__ == "__main__":
    main()
