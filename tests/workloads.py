"""Seeded synthetic point clouds for the parity tests and the benchmark.

The generator ``G`` draws, from ``numpy.random.default_rng(seed)`` and in this
order: ``n_blobs`` centres uniform in the centre box, then each blob's points
(``normal(centre, std)``; the first ``(n - n_noise) % n_blobs`` blobs get one
extra point), then ``n_noise = round(noise_frac * n)`` uniform noise points in
the noise box; it concatenates them, casts to float32 and shuffles the rows.

Named workloads:
  blobs10k   10,000 pts: 8 blobs std 0.5 in [-10,10]^2 + 5% noise in [-12,12]^2,
         eps 0.3, min_samples 10, seed 0
  lattice1m   1,000,000 pts: 100 blobs std 1.0 in [0,400]^2 + 10% noise, seed 1,
         then the 256x256 lattice {(0.5 i, 0.5 j)} appended in i-major order
         (N = 1,065,536); eps 0.5, min_samples 10
  blobs10m  10,000,000 pts: 1,000 blobs std 0.5 in [0,1265]^2 + 5% noise in
         [-126.5,1391.5]^2, seed 2; eps 0.5, min_samples 10
  skewed10m  10,000,000 pts, seed 3: 16 centres then 400 centres uniform in
         [0,2000]^2; 6M points in the 16 blobs (std 0.05), 3M in the 400 blobs
         (std 1.0), 1M uniform in [0,2000]^2; eps 0.25, min_samples 20
  clouds2048   2,048 clouds (master seed 4): N_i = floor(exp(U(ln 200, ln 5e5))),
         cloud i = G(N_i, max(1, round(8 N_i / 1e4)), [-L/2, L/2], 0.5, 0.05,
         [-0.6 L, 0.6 L], 1000 + i) with L = 20 sqrt(N_i / 1e4);
         eps_i = f32(U[0.2, 1.0]), min_samples_i from {5, 10, 20}
  serve    open-loop Poisson stream per rate (seed [5, rate]): N_i from the
         mixture 70% log-U[1k,50k], 25% log-U[50k,1M], 5% log-U[1M,10M],
         eps_i and min_samples_i as clouds2048, clouds at blobs10k's density;
         request i of the stream at `rate` has its own cloud,
         cloud_at_density(N_i, 100000 * rate + i) (serve_request_cloud)
"""

from __future__ import annotations

import numpy as np


def G(n, n_blobs, center_box, std, noise_frac, noise_box, seed) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n_noise = int(round(noise_frac * n))
    lo, hi = center_box
    centers = rng.uniform(lo, hi, (n_blobs, 2))
    n_blob_pts = n - n_noise
    base, extra = divmod(n_blob_pts, n_blobs)
    parts = []
    for b in range(n_blobs):
        cnt = base + (1 if b < extra else 0)
        parts.append(rng.normal(centers[b], std, (cnt, 2)))
    nlo, nhi = noise_box
    parts.append(rng.uniform(nlo, nhi, (n_noise, 2)))
    X = np.concatenate(parts).astype(np.float32)
    return _shuffle_rows(rng, X)


def _shuffle_rows(rng, X: np.ndarray) -> np.ndarray:
    """``rng.shuffle(X)`` for a 2-D array, 20x faster: Generator.shuffle draws
    the same swap sequence for a 1-D index array as for the rows (checked in
    tests/test_oracle.py), so the rows are gathered once by that index."""
    idx = np.arange(len(X), dtype=np.int64)
    rng.shuffle(idx)
    return X[idx]


def lattice(side: int = 256, spacing: float = 0.5) -> np.ndarray:
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    return np.stack([i.ravel() * spacing, j.ravel() * spacing], axis=1).astype(np.float32)


def blobs10k():
    return G(10_000, 8, (-10, 10), 0.5, 0.05, (-12, 12), 0), 0.3, 10


def lattice1m():
    X = G(1_000_000, 100, (0, 400), 1.0, 0.10, (0, 400), 1)
    return np.concatenate([X, lattice()]), 0.5, 10


def blobs10m():
    return G(10_000_000, 1000, (0, 1265), 0.5, 0.05, (-126.5, 1391.5), 2), 0.5, 10


def skewed10m():
    rng = np.random.default_rng(3)
    c16 = rng.uniform(0, 2000, (16, 2))
    c400 = rng.uniform(0, 2000, (400, 2))
    parts = [rng.normal(c, 0.05, (6_000_000 // 16, 2)) for c in c16]
    parts += [rng.normal(c, 1.0, (3_000_000 // 400, 2)) for c in c400]
    parts.append(rng.uniform(0, 2000, (1_000_000, 2)))
    X = np.concatenate(parts).astype(np.float32)
    return _shuffle_rows(rng, X), 0.25, 20


def clouds2048_params(n_clouds: int = 2048, seed: int = 4):
    """Per-cloud (N_i, eps_i, min_samples_i) of the batched workload."""
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp(rng.uniform(np.log(200), np.log(500_000), n_clouds))).astype(np.int64)
    eps = rng.uniform(0.2, 1.0, n_clouds).astype(np.float32)
    ms = rng.choice(np.array([5, 10, 20]), n_clouds)
    return ns, eps, ms


def clouds2048_cloud(i: int, n_i: int) -> np.ndarray:
    return cloud_at_density(n_i, 1000 + i)


def cloud_at_density(n: int, seed: int) -> np.ndarray:
    """A cloud of n points at blobs10k's density (the generator the batched
    and served workloads draw from)."""
    L = 20.0 * np.sqrt(n / 1e4)
    nb = max(1, int(round(8 * n / 1e4)))
    return G(int(n), nb, (-L / 2, L / 2), 0.5, 0.05, (-0.6 * L, 0.6 * L), seed)


def serve_stream(rate: float, count: int, seed: int = 5):
    """Open-loop Poisson request stream: arrival times (s), N_i from the
    mixture 70% log-U[1k, 50k], 25% log-U[50k, 1M], 5% log-U[1M, 10M], eps_i
    = f32(U[0.2, 1.0]), min_samples_i from {5, 10, 20}."""
    rng = np.random.default_rng([seed, int(rate)])
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    u = rng.uniform(0, 1, count)
    lo = np.where(u < 0.70, 1e3, np.where(u < 0.95, 5e4, 1e6))
    hi = np.where(u < 0.70, 5e4, np.where(u < 0.95, 1e6, 1e7))
    ns = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi)))).astype(np.int64)
    eps = rng.uniform(0.2, 1.0, count).astype(np.float32)
    ms = rng.choice(np.array([5, 10, 20]), count)
    return arrivals, ns, eps, ms


# The pinned served sets: the first SERVE_PINNED_COUNT requests of the streams
# at these rates, each with its own cloud (tests/golden/streams.json).
SERVE_PINNED_RATES = (1000, 5000)
SERVE_PINNED_COUNT = 300


def serve_request_cloud(rate: int, i: int, n: int) -> np.ndarray:
    """The points of request i of the stream at `rate` req/s (n = its N_i)."""
    return cloud_at_density(int(n), 100_000 * int(rate) + int(i))


WORKLOADS = {"blobs10k": blobs10k, "lattice1m": lattice1m, "blobs10m": blobs10m, "skewed10m": skewed10m}


def cell_keys_np(X: np.ndarray, eps: float):
    """(kx, ky) int64 arrays by the fp32 cell rule floor(f32(x) * inv_r),
    inv_r = f32(sqrt 2) / f32(eps); finite in-range inputs only."""
    inv_r = np.float32(np.sqrt(2.0)) / np.float32(eps)
    kx = np.floor(X[:, 0].astype(np.float32) * inv_r).astype(np.int64)
    ky = np.floor(X[:, 1].astype(np.float32) * inv_r).astype(np.int64)
    return kx, ky
