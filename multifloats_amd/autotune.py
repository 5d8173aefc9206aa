"""Tile-shape autotuner for the grouped (batched) DD GEMM.

The grouped kernel has six GEMM tile shapes (variant: outputs per lane ->
tile): 0: 2x2 -> 32x32, 1: 4x4 -> 64x64, 2: 8x1 -> 128x16, 3: 1x8 -> 16x128,
4: 4x2 -> 64x32, 5: 2x4 -> 32x64. The tuner times the grouped launch once per
variant on a synthetic batch whose m and n are multiples of 128 (no padding
for any shape; k log-uniform in [8, 512]) and records each variant's rate in
DD-MAC/s. plan_variants() then gives every problem the variant with the least
estimated time, padded MACs / rate: small or skinny problems get the shapes
that waste the fewest lanes. The table is JSON
(multifloats_amd/tuning/grouped.json); without it the kernel's built-in rule
(least padded work with fixed per-shape weights) applies.

    python -m multifloats_amd.autotune      # needs a GPU; rewrites the table
"""
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "grouped.json")
TILES = [(32, 32), (64, 64), (128, 16), (16, 128), (64, 32), (32, 64)]   # (rows, cols) per variant
VARIANTS = list(range(len(TILES)))


def load_table(path=TABLE):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t if "variant_rates" in t else None


def padded_macs(m, k, n, v):
    bm, bn = TILES[v]
    return (-(-m // bm) * bm) * (-(-n // bn) * bn) * max(k, 1)


def plan_variants(sizes, table=None):
    """Tile variant per GEMM (sizes: (ngemm, 3) array of m, k, n)."""
    table = table if table is not None else load_table()
    if table is None:
        return [-1] * len(sizes)
    rate = table["variant_rates"]
    out = []
    for m, k, n in sizes:
        m, k, n = int(m), int(k), int(n)
        out.append(int(np.argmin([padded_macs(m, k, n, v) / rate[v] for v in VARIANTS])))
    return out


def tune(torch, mf, problems=400, reps=3, seed=11, path=TABLE):
    from . import datagen
    rng = np.random.default_rng(seed)
    stream = torch.cuda.current_stream().cuda_stream
    mn = rng.choice([128, 256, 384], size=(problems, 2))
    ks = np.clip(np.round(np.exp(rng.uniform(np.log(8), np.log(512), problems))), 8, 512).astype(np.int64)
    tot = int(sum(m * k + k * n + m * n for (m, n), k in zip(mn, ks)))
    buf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, buf, seed)
    base = buf.data_ptr()
    off = 0
    gemms = []
    for (m, n), k in zip(mn, ks):
        m, n, k = int(m), int(n), int(k)
        gemms.append((base + 16 * off, base + 16 * (off + m * k), base + 16 * (off + m * k + k * n), m, k, n, 8))
        off += m * k + k * n + m * n
    macs = float(sum(int(m) * int(k) * int(n) for (m, n), k in zip(mn, ks)))
    rates = []
    for v in VARIANTS:
        plan = mf.GroupedPlan(gemms, [], [v] * len(gemms))
        plan.run_graph(stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.run_graph(stream)
        torch.cuda.synchronize()
        rates.append(macs / ((time.perf_counter() - t0) / reps))
        plan.close()
    del buf
    result = {"tiles": TILES, "variant_rates": [round(r, 1) for r in rates],
              "dd_gmacs": [round(r / 1e9, 1) for r in rates],
              "batch": "%d problems, m, n in {128, 256, 384}, k log-U[8, 512]" % problems}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    return result


if __name__ == "__main__":
    import torch
    import multifloats_amd as mf
    print(json.dumps(tune(torch, mf), indent=1))
