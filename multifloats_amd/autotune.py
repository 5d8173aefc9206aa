"""Tile-shape autotuner for the grouped (batched) DD GEMM.

The grouped kernel has two GEMM tile variants (0: 32x32 with 2x2 outputs per
lane, 1: 64x64 with 4x4). Problems are binned into size classes by
max(m, n); for every class the tuner times the grouped launch on a synthetic
batch drawn from that class with each variant and keeps the faster one. The
table is stored as JSON (multifloats_amd/tuning/grouped.json) and used by
plan_variants(); without a table the kernel's built-in rule applies.

    python -m multifloats_amd.autotune      # needs a GPU; rewrites the table
"""
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "grouped.json")
CLASS_EDGES = [24, 48, 96, 192]          # class c: edges[c-1] < max(m, n) <= edges[c]
VARIANTS = [0, 1]


def size_class(m, n):
    x = max(m, n)
    for c, e in enumerate(CLASS_EDGES):
        if x <= e:
            return c
    return len(CLASS_EDGES)


def load_table(path=TABLE):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def plan_variants(sizes, table=None):
    """Tile variant per GEMM (sizes: (ngemm, 3) array of m, k, n)."""
    table = table if table is not None else load_table()
    if table is None:
        return [-1] * len(sizes)
    best = table["variant_by_class"]
    return [int(best[size_class(int(m), int(n))]) for m, k, n in sizes]


def tune(torch, mf, problems_per_class=300, reps=3, seed=11, path=TABLE):
    from . import datagen
    rng = np.random.default_rng(seed)
    edges = [8] + CLASS_EDGES + [512]
    stream = torch.cuda.current_stream().cuda_stream
    result = {"class_edges": CLASS_EDGES, "variant_by_class": [], "measured": []}
    for c in range(len(edges) - 1):
        lo, hi = edges[c] + (1 if c else 0), edges[c + 1]
        mn = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), (problems_per_class, 2)))).astype(np.int64)
        mn = np.clip(mn, lo, hi)
        ks = np.clip(np.round(np.exp(rng.uniform(np.log(8), np.log(512), problems_per_class))), 8, 512).astype(np.int64)
        tot = int(sum(m * k + k * n + m * n for (m, n), k in zip(mn, ks)))
        buf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
        datagen.fill_config0_torch(torch, buf, seed + c)
        base = buf.data_ptr()
        off = 0
        gemms = []
        for (m, n), k in zip(mn, ks):
            a = base + 16 * off
            off += m * k
            b = base + 16 * off
            off += k * n
            cc = base + 16 * off
            off += m * n
            gemms.append((a, b, cc, int(m), int(k), int(n), 8))
        macs = float(sum(int(m) * int(k) * int(n) for (m, n), k in zip(mn, ks)))
        times = []
        for v in VARIANTS:
            plan = mf.GroupedPlan(gemms, [], [v] * len(gemms))
            plan.run_graph(stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                plan.run_graph(stream)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) / reps)
            plan.close()
        best = VARIANTS[int(np.argmin(times))]
        result["variant_by_class"].append(best)
        result["measured"].append({"class": c, "max_mn": [lo, hi],
                                   "dd_gmacs": [round(macs / t / 1e9, 1) for t in times]})
        del buf
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    return result


if __name__ == "__main__":
    import torch
    import multifloats_amd as mf
    print(json.dumps(tune(torch, mf), indent=1))
