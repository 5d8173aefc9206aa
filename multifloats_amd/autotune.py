"""Tile-shape autotuner for the grouped (batched) DD GEMM, per size class.

The grouped kernels have eight GEMM tile shapes (variant: outputs per lane ->
tile): 0: 2x2 -> 32x32, 1: 4x4 -> 64x64, 2: 8x1 -> 128x16, 3: 1x8 -> 16x128,
4: 4x2 -> 64x32, 5: 2x4 -> 32x64, 6: 1x1 -> 16x16, 7: 3x3 -> 48x48.

What a shape costs on a problem is its padded MACs (exact: the tile grid over
m x n, times k) divided by the shape's rate, and the rate depends on k: each
tile pays a fixed cost (operand staging, barriers, the final renormalisation
and store) against its k steps of MACs. So the size classes are k buckets
(K_EDGES; the C4 mix draws m, n, k log-uniform in [8, 512]), and the tuner
times every shape on every class: a batch of problems with m = n = 384 (no
padding for any shape) and k log-uniform inside the class, in one grouped
launch replayed from a graph, rate = MACs / time. plan_variants() then gives
each GEMM the shape with the least padded MACs / rate of its class. The table
is JSON (multifloats_amd/tuning/grouped.json); without it the library's
built-in rule (least padded work with fixed per-shape weights) applies.

    python -m multifloats_amd.autotune      # needs a GPU; rewrites the table
"""
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "grouped.json")
TILES = [(32, 32), (64, 64), (128, 16), (16, 128), (64, 32), (32, 64), (16, 16), (48, 48)]  # (rows, cols)
VARIANTS = list(range(len(TILES)))
K_EDGES = [8, 24, 64, 160, 513]  # size classes: k in [K_EDGES[c], K_EDGES[c + 1])


def k_class(k):
    for c in range(len(K_EDGES) - 1):
        if k < K_EDGES[c + 1]:
            return c
    return len(K_EDGES) - 2


def load_table(path=TABLE):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    ok = "class_rates" in t and t.get("tiles") == [list(x) for x in TILES] and t.get("k_edges") == K_EDGES
    return t if ok else None


def padded_macs(m, k, n, v):
    bm, bn = TILES[v]
    return (-(-m // bm) * bm) * (-(-n // bn) * bn) * max(k, 1)


def plan_variants(sizes, table=None):
    """Tile variant per GEMM (sizes: (ngemm, 3) array of m, k, n)."""
    table = table if table is not None else load_table()
    if table is None:
        return [-1] * len(sizes)
    rates = table["class_rates"]
    out = []
    for m, k, n in sizes:
        m, k, n = int(m), int(k), int(n)
        r = rates[k_class(k)]
        out.append(int(np.argmin([padded_macs(m, k, n, v) / r[v] for v in VARIANTS])))
    return out


def tune(torch, mf, problems=120, reps=3, seed=11, path=TABLE):
    from . import datagen
    rng = np.random.default_rng(seed)
    stream = torch.cuda.current_stream().cuda_stream
    edge = 384  # a multiple of every tile edge
    rates = []
    for c in range(len(K_EDGES) - 1):
        lo, hi = K_EDGES[c], K_EDGES[c + 1] - 1
        ks = np.clip(np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), problems))), lo, hi).astype(np.int64)
        tot = int(sum(2 * edge * k + edge * edge for k in ks))
        buf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
        datagen.fill_config0_torch(torch, buf, seed + c)
        base, off, gemms = buf.data_ptr(), 0, []
        for k in ks:
            k = int(k)
            gemms.append((base + 16 * off, base + 16 * (off + edge * k), base + 16 * (off + 2 * edge * k),
                          edge, k, edge, 8))
            off += 2 * edge * k + edge * edge
        macs = float(edge * edge * ks.sum())
        row = []
        for v in VARIANTS:
            plan = mf.GroupedPlan(gemms, [], [v] * len(gemms))
            plan.run_graph(stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                plan.run_graph(stream)
            torch.cuda.synchronize()
            row.append(round(macs / ((time.perf_counter() - t0) / reps), 1))
            plan.close()
        rates.append(row)
        del buf
    result = {"tiles": TILES, "k_edges": K_EDGES, "class_rates": rates,
              "dd_gmacs": [[round(r / 1e9, 1) for r in row] for row in rates],
              "batch": "per class: %d problems, m = n = %d, k log-uniform in the class" % (problems, edge)}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    return result


if __name__ == "__main__":
    import torch
    import multifloats_amd as mf
    print(json.dumps(tune(torch, mf)["dd_gmacs"], indent=1))
