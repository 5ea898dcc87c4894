"""Autotuner for the launch configuration of the gfx950 pipeline.

No counterpart in the reference (its launch shapes are fixed).  A
configuration is (sort_items, radix_bits, cell_blocks, pair_blocks):

  sort_items   points per thread of a radix-sort tile (tile = 512 x items):
               8, 12 or 16
  radix_bits   widest radix digit the sort may use when it saves no pass
  cell_blocks  workgroups of the per-cell kernels (core marking, border
               assignment): how many cells each workgroup walks
  pair_blocks  workgroups of the cell-pair union kernels

Every configuration gives the same labels (the tuner checks it on every
candidate); they differ in speed only.  For each size bucket the tuner times
candidates on a synthetic cloud at constant density by coordinate descent
(sort shape first, then the two grid sizes), keeps the fastest, and persists
the table as JSON.  ``load()`` installs a saved table for this process; the
package loads ``tuning_gfx950.json`` beside this file at import when present.

    python -m dbscan_torch.autotune          # tune and save
"""

from __future__ import annotations

import json
import math
import os
import sys
import time
from typing import Dict, List, Optional, Sequence

import torch

from . import _C

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")

# representative size of each bucket; a bucket covers sizes up to the
# geometric midpoint with the next one
BUCKETS = (10_000, 100_000, 1_000_000, 10_000_000)
SORT_SHAPES = [(items, bits) for items in (8, 12, 16) for bits in (7, 8, 9)]
CELL_BLOCKS = (1024, 2048, 4096, 8192, 16384, 32768)
PAIR_BLOCKS = (2048, 4096, 8192, 16384, 32768, 65536)
INT32_MAX = 2**31 - 1


def synthetic_cloud(n: int, seed: int) -> torch.Tensor:
    """n float32 points at constant density: max(1, round(8 n / 1e4))
    Gaussian blobs (std 0.5) in a square of side L = 20 sqrt(n / 1e4), plus
    5% uniform noise over 1.2 L."""
    g = torch.Generator().manual_seed(seed)
    L = 20.0 * math.sqrt(n / 1e4)
    nb = max(1, int(round(8 * n / 1e4)))
    n_noise = int(round(0.05 * n))
    centers = (torch.rand((nb, 2), generator=g, dtype=torch.float64) - 0.5) * L
    which = torch.randint(0, nb, (n - n_noise,), generator=g)
    blobs = centers[which] + 0.5 * torch.randn((n - n_noise, 2), generator=g, dtype=torch.float64)
    noise = (torch.rand((n_noise, 2), generator=g, dtype=torch.float64) - 0.5) * (1.2 * L)
    X = torch.cat([blobs, noise]).to(torch.float32)
    return X[torch.randperm(n, generator=g)]


def _time(X, eps, ms, cfg, steps) -> tuple:
    labels, _, _, total = _C.dbscan2d_benchmark(X, eps, ms, steps, list(cfg))
    return total / steps, labels


def tune_bucket(n: int, eps: float = 0.5, min_samples: int = 10, steps: int = 5,
                device="cuda", log=None) -> dict:
    """Coordinate descent over the candidates for one cloud of n points."""
    X = synthetic_cloud(n, 12345 + n).to(device)
    base = list(_C.default_config(n))
    _time(X, eps, min_samples, base, 2)  # warm-up
    best_ms, ref = _time(X, eps, min_samples, base, steps)
    best = base
    tried: Dict[str, float] = {str(tuple(base)): round(best_ms, 4)}

    def consider(cfg):
        nonlocal best, best_ms
        key = str(tuple(cfg))
        if key in tried:
            return
        _time(X, eps, min_samples, cfg, 1)
        t, labels = _time(X, eps, min_samples, cfg, steps)
        if not torch.equal(labels, ref):
            raise AssertionError(f"autotune: configuration {cfg} changed the labels at n={n}")
        tried[key] = round(t, 4)
        if t < best_ms:
            best, best_ms = list(cfg), t

    for items, bits in SORT_SHAPES:
        consider([items, bits, best[2], best[3]])
    for cb in CELL_BLOCKS:
        consider([best[0], best[1], cb, best[3]])
    for pb in PAIR_BLOCKS:
        consider([best[0], best[1], best[2], pb])
    if log:
        log(f"[autotune] n={n}: best {best} {best_ms:.4f} ms (default {tried[str(tuple(base))]} ms)")
    return {"n": n, "config": best, "ms": round(best_ms, 4),
            "default_ms": tried[str(tuple(base))], "tried": tried}


def autotune(buckets: Sequence[int] = BUCKETS, steps: int = 5, save: Optional[str] = TABLE_PATH,
             install: bool = True, log=None) -> dict:
    """Tunes every bucket, installs the table for this process and saves it."""
    results = [tune_bucket(n, steps=steps, log=log) for n in buckets]
    rows = []
    for i, r in enumerate(results):
        hi = (INT32_MAX if i + 1 == len(results)
              else int(math.sqrt(results[i]["n"] * results[i + 1]["n"])))
        rows.append({"max_n": hi, "config": r["config"]})
    table = {"device": torch.cuda.get_device_name(0), "torch": torch.__version__,
             "created": time.strftime("%Y-%m-%d %H:%M:%S"), "rows": rows, "buckets": results}
    if install:
        install_rows(rows)
    if save:
        with open(save, "w") as f:
            json.dump(table, f, indent=1)
    return table


def install_rows(rows: List[dict]) -> None:
    _C.set_tuning_table([[int(r["max_n"])] + [int(v) for v in r["config"]] for r in rows])


def load(path: str = TABLE_PATH) -> bool:
    """Installs a saved table for this process; False if there is none."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        table = json.load(f)
    install_rows(table["rows"])
    return True


if __name__ == "__main__":
    out = autotune(log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps({"rows": out["rows"]}))
