#!/usr/bin/env python3
"""Benchmark of the DD hot path on one MI355X.

Default (the headline): float64x2 GEMM C = A*B at m = n = k = 4096, entries
s*2^U(-20,20)*u with lo per the config-0 rule, inputs and output resident in
HBM, in two forms on the same inputs:
  ozaki  the INT8-MFMA split-precision GEMM (slices chosen from the data), whole
         call per step (exponent scan, residues, products, reconstruction);
  plain  the reference-order FP64 GEMM (bit-identical to the reference),
         renorm interval 8.
Prints ONE JSON line: value = Ozaki DD-GFLOP/s (2*m*n*k per step) when its
measured error on a sampled block is within 2^-100 (|A||B|) elementwise, with
the plain form, the error figures and the speed-up in the same line.

  python bench.py [--steps K] [--warmup W] [--workload gemm|cgemm|blas|trsm|batched|elementwise|reduce]
                  [--no-cpu] [--variant V]

Timing: torch.cuda events recorded on the stream the kernels are launched on
(torch's current stream, whose handle is passed to the C ABI), bracketed by
device synchronisation. The CPU baseline is the reference's own compiled
matmul (oracle/_ref) when present, else the oracle's C restatement, timed on
one host core over a bounded slice (all of A times the first columns of B)
of the same workload; its outputs are also compared bit-for-bit with the GPU's.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_FLOPS_PER_DDMAC = 15.75  # 2 MUL + 2 FMA(x2) + 9 ADD + 6/8 renorm at ri = 8
INT8_PEAK_TOPS = 5000.0      # MI355X dense INT8 MFMA

# HBM-side bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (KiB per dispatch) on the 4096^3 workload, scaled by what the same counters
# read for a 1 GiB device copy: FETCH_SIZE x 2.0 (it reads 524,300 KiB for
# 1 GiB fetched), WRITE_SIZE x 1.0 (1,048,576 KiB). Summary and commands:
# profiles/pmc_traffic_gemm_4096.md. Keyed by (kernel, size, slices).
PMC_TRAFFIC = {
    ("k_ozaki", 4096, 14): 2765231.0 * 2.0 * 1024 + 286820.0 * 1024,
    ("k_oz2_gemm", 4096, 14): 1682594.0 * 2.0 * 1024 + 599483.0 * 1024,
    ("k_gemm_plain", 4096, 0): 3144165.19 * 2.0 * 1024 + 286727.06 * 1024,
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="gemm", choices=["gemm", "cgemm", "blas", "trsm", "batched", "elementwise", "reduce"])
    p.add_argument("--retune", action="store_true", help="batched: rerun the tile autotuner first")
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--spread", type=float, default=20.0)
    p.add_argument("--slices", type=int, default=0, help="0 = chosen from the data")
    p.add_argument("--form", default="both", choices=["both", "plain", "ozaki"])
    p.add_argument("--err-sample", type=int, default=256, help="error on an N x N block of sampled outputs")
    p.add_argument("--variant", type=int, default=-1)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-cols", type=int, default=512, help="CPU baseline: all of A times this many columns of B")
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


def cpu_gemm_baseline(A, B, C_gpu, cols, ri):
    """Time the reference matmul on one core over all of A times B[:, :cols]."""
    from tests import _native as nt  # test infrastructure, timed as the baseline
    from multifloats_amd import datagen
    m, k = A.shape[:2]
    Bs = datagen.as_colmajor(B[:, :cols])
    A = datagen.as_colmajor(A)
    C = datagen.as_colmajor(np.zeros((m, cols, 2)))
    kind = "reference"
    if os.path.exists(nt.REF):
        L = ctypes.CDLL(nt.REF)
        fn = L.matmuldd_mm
    else:
        kind = "port"
        fn = nt.oracle().or_matmul_mm
    fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64] * 4
    t0 = time.perf_counter()
    fn(A.ctypes.data, Bs.ctypes.data, C.ctypes.data, m, k, cols, ri)
    dt = time.perf_counter() - t0
    same = bool(nt.bits_equal(C, C_gpu[:, :cols]).all())
    return dict(value=round(2.0 * m * k * cols / dt / 1e9, 4), unit="DD-GFLOP/s", cores=1, kind=kind,
                sample="C[:, :%d] = A(%dx%d) * B[:, :%d] on one host core, %.1f s; bit-identical to the GPU: %s"
                % (cols, m, k, cols, dt, same)), same


def sampled_error(A, B, C, rows, cols):
    from tests import _native as nt  # exact accumulator (test infrastructure)
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    err, mag = nt.gemm_err(A, B, C, rr.ravel(), cc.ravel(), threads=min(16, os.cpu_count() or 1))
    rel = err / mag
    return float(rel.max()), float(rel.mean())


def bench_gemm(args, torch, mf):
    from multifloats_amd import datagen
    m = n = k = args.size
    ri = 8
    S = args.slices
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
        achieved = FP64_FLOPS_PER_DDMAC * m * n * k / (ms * 1e-3) / 1e12
        res["plain"] = dict(ms=ms, value=flops / (ms * 1e-3) / 1e9, err_max=emax, err_mean=emean,
                            roofline={"bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                                      "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                                      "traffic": PMC_TRAFFIC.get(("k_gemm_plain", args.size, 0))
                                      if args.variant < 0 else None,
                                      "note": "k_gemm_plain, FP64 VALU: %.2f FP64 flops per DD-MAC (FMA = 2)"
                                              % FP64_FLOPS_PER_DDMAC})
        if not args.no_cpu:
            res["cpu"], _ = cpu_gemm_baseline(A, B, C, args.cpu_cols, ri)
        del C

    if args.form in ("both", "ozaki"):
        wb = mf.ozaki_workspace_bytes(m, k, n, S)
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")

        def step_oz():
            mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, S, work.data_ptr(), wb, stream)

        def step_mult():
            mf.dev_ozaki_multiply(dC.data_ptr(), m, k, n, S, work.data_ptr(), wb, stream)

        def step_products():
            mf.dev_ozaki_products(m, k, n, S, work.data_ptr(), wb, stream)
        ms = timed(torch, step_oz, args.steps, args.warmup)
        S = mf.ozaki_slices_used()
        if S == 0:
            raise SystemExit("bench: the slice rule chose the reference-order GEMM for this input")
        mf.dev_ozaki_prepare(dA.data_ptr(), dB.data_ptr(), m, k, n, S, work.data_ptr(), wb, stream)
        ms_mult = timed(torch, step_mult, args.steps, 1)
        N = mf.ozaki_moduli(m, k, n, S)
        if N:
            # CRT form: the dominant kernel is k_oz2_gemm, one launch of the N
            # residue INT8 GEMMs, timed alone through the products entry point
            # (the multiply stage adds the reconstruction, k_oz2_crt)
            ms_kern = timed(torch, step_products, args.steps, 1)
            mf.dev_ozaki_multiply(dC.data_ptr(), m, k, n, S, work.data_ptr(), wb, stream)
            P, kern = N, "k_oz2_gemm"
            note = ("k_oz2_gemm: %d residue INT8 GEMMs x 2mnk int8 ops per launch (MFMA), timed alone; the "
                    "reconstruction k_oz2_crt takes the rest of ms_multiply" % N)
        else:
            ms_kern = ms_mult
            P, kern = S * (S + 1) // 2, "k_ozaki"
            note = "k_ozaki INT8 MFMA: %d slice pairs x 2mnk int8 ops per launch" % P
        C = datagen.unflat(dC.cpu().numpy(), m, n)
        emax, emean = sampled_error(A, B, C, rows, cols)
        tops = P * 2.0 * m * n * k / (ms_kern * 1e-3) / 1e12
        res["ozaki"] = dict(ms=ms, ms_multiply=ms_mult, ms_kernel=ms_kern, value=flops / (ms * 1e-3) / 1e9,
                            err_max=emax,
                            err_mean=emean, slices=S, form="crt" if N else "triangular",
                            int8_products=P,
                            roofline={"bound": "mfma", "achieved": round(tops, 2), "peak": INT8_PEAK_TOPS,
                                      "unit": "TOP/s", "frac": round(tops / INT8_PEAK_TOPS, 4),
                                      "traffic": PMC_TRAFFIC.get((kern, args.size, S)), "note": note})
        del C

    head = res.get("ozaki") if args.form != "plain" else res["plain"]
    ok = head["err_max"] <= 2.0 ** -100 if args.form != "plain" else True
    out = {
        "metric": "DD-GEMM DD-GFLOP/s at 4096^3: Ozaki-INT8 vs plain FP64, err<=2^-100*|A||B|",
        "value": round(head["value"], 2) if ok else None, "unit": "DD-GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8-mfma+f64" if args.form != "plain" else "f64", "data": "synthetic",
        "config": {"workload": "float64x2 GEMM m=n=k=%d, entries s*2^U(-%g,%g)*u, lo per config-0; %s"
                               % (m, args.spread, args.spread,
                                  "Ozaki INT8 (slices chosen from the data: %d) and plain reference-order "
                                  "(renorm 8) forms" % S
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
            out[form]["roofline"] = r["roofline"]
    if "ozaki" in res and "plain" in res:
        out["speedup_ozaki_vs_plain"] = round(res["plain"]["ms"] / res["ozaki"]["ms"], 3)
    if "cpu" in res:
        out["cpu_baseline"] = res["cpu"]
    return out


def bench_cgemm(args, torch, mf):
    """complex64x2 GEMM m = n = k = 2048 (4 real chains per output, fused)."""
    from multifloats_amd import datagen
    m = n = k = 2048 if args.size == 4096 else args.size
    rng = np.random.default_rng(3)

    def cm(r, c):
        re, im = datagen.matrix(rng, r, c, args.spread), datagen.matrix(rng, r, c, args.spread)
        return np.ascontiguousarray(np.concatenate([re, im], axis=2).transpose(1, 0, 2)).reshape(-1, 4)
    dA = torch.from_numpy(cm(m, k)).cuda()
    dB = torch.from_numpy(cm(k, n)).cuda()
    dC = torch.empty((m * n, 4), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    ms = timed(torch, lambda: mf.dev_cmatmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8, stream),
               args.steps, args.warmup)
    flops = 8.0 * m * n * k
    achieved = 4 * FP64_FLOPS_PER_DDMAC * m * n * k / (ms * 1e-3) / 1e12
    return {"metric": "complex DD-GEMM DD-GFLOP/s", "value": round(flops / (ms * 1e-3) / 1e9, 2),
            "unit": "DD-GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 4), "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "complex64x2 GEMM m=n=k=%d, spread 2^+-%g, renorm 8, bit-identical to the "
                                   "reference's four-GEMM composition" % (m, args.spread)},
            "roofline": {"bound": "fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None}}


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


def bench_batched(args, torch, mf):
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
    return {"metric": "batched DD GEMMs + dots, aggregate DD-GFLOP/s", "value": round(dd_flops / (ms * 1e-3) / 1e9, 2),
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


def bench_elementwise(args, torch, mf):
    from multifloats_amd import datagen
    n = 1 << 26
    rng = np.random.default_rng(7)
    stream = torch.cuda.current_stream().cuda_stream
    res = {}
    a = torch.from_numpy(datagen.config0(rng, n)).cuda()
    b = torch.from_numpy(datagen.config0(rng, n)).cuda()
    c = torch.empty_like(a)
    for nm, op in (("add", mf.MFHIP_ADD), ("mul", mf.MFHIP_MUL), ("div", mf.MFHIP_DIV)):
        ms = timed(torch, lambda: mf.dev_binary(op, a.data_ptr(), b.data_ptr(), c.data_ptr(), n, stream),
                   args.steps, args.warmup)
        res[nm] = dict(ms=round(ms, 4), GBs=round(48.0 * n / ms / 1e6, 1))
    a.abs_()
    ms = timed(torch, lambda: mf.dev_unary(mf.MFHIP_SQRT, a.data_ptr(), c.data_ptr(), n, stream), args.steps, args.warmup)
    res["sqrt"] = dict(ms=round(ms, 4), GBs=round(32.0 * n / ms / 1e6, 1))
    # two-argument functions on x in (0, 1], y in [-1, 1]: pow takes its 2^(y log2 x) path
    for nm, op in (("pow", mf.MFHIP_POW), ("hypot", mf.MFHIP_HYPOT), ("fmax", mf.MFHIP_FMAX), ("atan2", mf.MFHIP_ATAN2)):
        ms = timed(torch, lambda: mf.dev_binary(op, a.data_ptr(), b.data_ptr(), c.data_ptr(), n, stream),
                   args.steps, args.warmup)
        res[nm] = dict(ms=round(ms, 4), GBs=round(48.0 * n / ms / 1e6, 1), Gelem_s=round(n / ms / 1e6, 2))
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
    ops = {nm: getattr(mf, "MFHIP_" + nm.upper()) for nm in gens}
    for nm, gen in gens.items():
        x = torch.from_numpy(gen(rng, n)).cuda()
        ms = timed(torch, lambda: mf.dev_unary(ops[nm], x.data_ptr(), c.data_ptr(), n, stream), args.steps, args.warmup)
        res[nm] = dict(ms=round(ms, 4), GBs=round(32.0 * n / ms / 1e6, 1), Gelem_s=round(n / ms / 1e6, 2))
        del x
    best = max(v["GBs"] for k_, v in res.items() if k_ in ("add", "mul", "div", "sqrt"))
    return {"metric": "elementwise DD arrays, N=2^26", "value": best, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "elementwise float64x2 N=2^26"},
            "roofline": {"bound": "hbm", "achieved": best, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(best / HBM_PEAK_GBS, 4), "traffic": None}, "ops": res}


def bench_reduce(args, torch, mf):
    from multifloats_amd import datagen
    n = 1 << 26
    rng = np.random.default_rng(8)
    a_np, b_np = datagen.cancel_pairs(rng, n)
    a = torch.from_numpy(a_np).cuda()
    b = torch.from_numpy(b_np).cuda()
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    work = torch.empty(mf.reduce_workspace_bytes(n) // 8 + 2, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    res = {}
    for nm, kind, bytes_per in (("dot", mf.MFHIP_DOT, 32), ("sum", mf.MFHIP_SUM, 16), ("norm2", mf.MFHIP_NORM2, 16),
                                ("product", mf.MFHIP_PRODUCT, 16), ("maxval", mf.MFHIP_MAXVAL, 16),
                                ("minval", mf.MFHIP_MINVAL, 16)):
        ms = timed(torch, lambda: mf.dev_reduce(kind, a.data_ptr(), b.data_ptr(), n, out.data_ptr(),
                                                work.data_ptr(), 8, stream), args.steps, args.warmup)
        res[nm] = dict(ms=round(ms, 4), GBs=round(bytes_per * n / ms / 1e6, 1))
    best = res["dot"]["GBs"]
    return {"metric": "DD dot_product N=2^26", "value": best, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "dot/sum/norm2/product/maxval/minval float64x2 N=2^26"},
            "roofline": {"bound": "hbm", "achieved": best, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(best / HBM_PEAK_GBS, 4), "traffic": None}, "ops": res}


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
    fn = {"gemm": bench_gemm, "cgemm": bench_cgemm, "blas": bench_blas, "trsm": bench_trsm, "batched": bench_batched,
          "elementwise": bench_elementwise, "reduce": bench_reduce}[args.workload]
    print(json.dumps(fn(args, torch, mf)), flush=True)


if __name__ == "__main__":
    main()
