"""Small constructed inputs whose reference labels are frozen in cases.npz.

Each case isolates one rule of the reference CPU path
(csrc/cpu/dbscan_grid.hpp) where a plausible mistake changes the labels:

  border_tie_offset   a border point at exactly eps from cores of two clusters;
                      the reference picks the lower window offset (:437), the
                      reference CUDA path's (d2, original index) rule would not
  signed_kx_order     clusters on both sides of kx = 0: numbering follows the
                      signed key (:176-178), not the unsigned one
  ky_straddle         clusters in one kx column at ky < 0 and ky >= 0: the
                      signed int64 key orders ky as uint32, so ky >= 0 first
  ring2_exact_eps     ring-2 pairs at exactly eps and one ulp either side
                      (the construction of tests/test_eps_boundary.py:56-93)
  ring2_corner        the extreme diagonal corner of a ring-2 offset
  same_cell_far       points whose cell coordinates overflow int32 share the
                      INT32_MIN column and count as neighbours wholesale (:295-297)
  nan_inf             NaN and infinite coordinates among ordinary clusters
  duplicates          many coincident points
  eps_denormal        eps whose float32 square underflows; inv_r overflows
  eps_huge            eps whose float32 square overflows
  ms_one              min_samples = 1: every point is core
  lattice_ties        a lattice at spacing exactly eps (d2 == eps_sq ties)
"""

from __future__ import annotations

import numpy as np


def _f32(v) -> np.float32:
    return np.float32(v)


def _boundary_float(eps: float, cell: int, side: str) -> float:
    """Largest ("high") or smallest ("low") float32 x with floor(x*inv_r) == cell."""
    inv_r = _f32(np.sqrt(_f32(2.0))) / _f32(eps)
    x = _f32(cell + (1 if side == "high" else 0)) / inv_r
    towards = _f32(-np.inf) if side == "high" else _f32(np.inf)
    while int(np.floor(x * inv_r)) != cell:
        x = np.nextafter(x, towards)
    while int(np.floor(np.nextafter(x, -towards) * inv_r)) == cell:
        x = np.nextafter(x, -towards)
    return float(x)


def _group(p, k):
    return [p] * k


def constructed_cases() -> dict:
    cases = {}
    ms = 5
    # Border tie: P at the origin, cores CA (-1, 0) and CB (1, 0) at d2 == 1 ==
    # eps^2.  CA is core through ms-1 points at (-1.6, 0) (within eps of CA,
    # beyond eps of P); CB likewise.  CB's cluster comes first in input order.
    pts = _group((1.0, 0.0), 1) + _group((1.6, 0.0), ms - 1)
    pts += [(0.0, 0.0)]
    pts += _group((-1.0, 0.0), 1) + _group((-1.6, 0.0), ms - 1)
    cases["border_tie_offset"] = (np.array(pts, np.float32), 1.0, ms)

    rng = np.random.default_rng(11)
    blobs = [rng.normal(c, 0.2, (60, 2)) for c in [(-7.0, 1.0), (6.0, -2.0), (-0.3, 4.0), (0.4, -4.0)]]
    X = np.concatenate(blobs).astype(np.float32)
    rng.shuffle(X)
    cases["signed_kx_order"] = (X, 0.5, 5)

    blobs = [rng.normal((0.1, c), 0.1, (50, 2)) for c in (-6.0, -2.5, 2.5, 6.0)]
    X = np.concatenate(blobs).astype(np.float32)
    rng.shuffle(X)
    cases["ky_straddle"] = (X, 0.5, 5)

    eps = 1.0
    just_over = float(np.nextafter(_f32(1.5), _f32(2.0)))
    just_under = float(np.nextafter(_f32(1.5), _f32(0.0)))
    pts = []
    for k, (a, b) in enumerate([((0.5, 0.5), (1.5, 0.5)), ((0.5, 0.5), (just_over, 0.5)),
                                ((0.5, 0.5), (just_under, 0.5))]):
        off = 10.0 * k  # separate the three pairs by far more than eps
        pts += _group((a[0] + off, a[1]), ms) + _group((b[0] + off, b[1]), ms)
    cases["ring2_exact_eps"] = (np.array(pts, np.float32), eps, ms)

    hi = _boundary_float(eps, 0, "high")
    lo = _boundary_float(eps, 2, "low")
    cases["ring2_corner"] = (np.array(_group((hi, hi), ms) + _group((lo, lo), ms), np.float32), eps, ms)

    pts = _group((1.0e9, 0.1), 3) + _group((-1.0e9, 0.1), 3) + _group((3.0e9, 0.2), 2)
    pts += [(0.0, 0.0), (0.1, 0.1), (0.2, 0.0)]
    cases["same_cell_far"] = (np.array(pts, np.float32), 0.5, 5)

    X = np.concatenate([rng.normal(c, 0.15, (40, 2)) for c in [(0, 0), (3, 3)]]).astype(np.float32)
    extra = np.array([[np.nan, 0.0], [0.0, np.nan], [np.nan, np.nan], [np.inf, 0.0],
                      [-np.inf, 0.0], [0.0, np.inf], [np.inf, np.inf], [np.nan, 0.05],
                      [np.inf, 0.05], [-np.inf, -np.inf]], np.float32)
    X = np.concatenate([X, extra])
    rng.shuffle(X)
    cases["nan_inf"] = (X, 0.5, 3)

    X = np.concatenate([np.repeat([[1.0, 1.0]], 30, 0), np.repeat([[1.2, 1.0]], 3, 0),
                        np.repeat([[5.0, 5.0]], 4, 0), rng.uniform(0, 6, (40, 2))]).astype(np.float32)
    rng.shuffle(X)
    cases["duplicates"] = (X, 0.3, 5)

    X = np.concatenate([np.zeros((4, 2)), rng.normal(0, 1, (20, 2))]).astype(np.float32)
    cases["eps_denormal"] = (X, 1e-40, 3)
    cases["eps_huge"] = (rng.normal(0, 1, (30, 2)).astype(np.float32), 1e30, 10)

    cases["ms_one"] = (rng.uniform(-3, 3, (300, 2)).astype(np.float32), 0.2, 1)

    i, j = np.meshgrid(np.arange(12), np.arange(12), indexing="ij")
    lat = np.stack([i.ravel() * 0.5 - 2.0, j.ravel() * 0.5 - 3.0], 1).astype(np.float32)
    lat = lat[(i.ravel() % 5 != 2)]  # cut the lattice into strips one gap apart
    cases["lattice_ties"] = (lat, 0.5, 5)
    return cases
