"""Pins the C oracle (oracle/dbscan_oracle.c) before anything is checked
against it.  CPU only.

* golden vectors written by the reference's own compiled CPU path
  (tests/golden/make_golden.py): blobs10k labels, constructed cases, label hashes;
* the compiled reference itself (oracle/_ref), where it is built;
* an O(N^2) brute force that states each reference rule directly
  (tests/bruteforce.py), on random and constructed inputs;
* closed forms and special cases (empty, ms = 1, ms > N, single point,
  narrowing of min_samples to int32);
* invariants: permutation equivariance and thread-count independence;
* sklearn: core and noise sets on blobs10k.
Each check is chosen so that a plausible mistake (a dropped own-cell term, a
wrong tiebreak, unsigned key order, an off-by-one window) fails it; see
test_planted_mistakes_are_caught.
"""

import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from tests import workloads as W
from tests.bruteforce import bruteforce_dbscan, canonical
from tests.golden.cases import constructed_cases


def sha(labels):
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden(golden_dir):
    with open(os.path.join(golden_dir, "golden.json")) as f:
        return json.load(f)


def test_blobs10k_matches_reference_golden(golden_dir, golden):
    X, eps, ms = W.blobs10k()
    labels = oracle.dbscan2d(X, eps, ms)
    ref = np.load(os.path.join(golden_dir, "blobs10k_labels.npy"), allow_pickle=False)
    assert np.array_equal(labels, ref)
    assert sha(labels) == golden["blobs10k"]["sha256"]


def test_lattice1m_matches_reference_hash(golden):
    X, eps, ms = W.lattice1m()
    labels = oracle.dbscan2d(X, eps, ms)
    g = golden["lattice1m"]
    assert (int(labels.max()) + 1, int((labels == -1).sum())) == (g["n_clusters"], g["n_noise"])
    assert sha(labels) == g["sha256"]


@pytest.mark.parametrize("name", sorted(constructed_cases()))
def test_constructed_cases_match_reference(golden_dir, name):
    X, eps, ms = constructed_cases()[name]
    z = np.load(os.path.join(golden_dir, "cases.npz"), allow_pickle=False)
    assert np.array_equal(z[f"x_{name}"], X, equal_nan=True)
    assert np.array_equal(oracle.dbscan2d(X, eps, ms), z[f"labels_{name}"])


def _finite_case(name):
    X, eps, ms = constructed_cases()[name]
    return np.isfinite(X).all() and np.abs(X).max() < 1e6 and 1e-6 < eps < 1e6


@pytest.mark.parametrize("name", [n for n in sorted(constructed_cases()) if _finite_case(n)])
def test_constructed_cases_match_bruteforce(name):
    X, eps, ms = constructed_cases()[name]
    assert np.array_equal(oracle.dbscan2d(X, eps, ms), bruteforce_dbscan(X, eps, ms))


def _random_inputs():
    rng = np.random.default_rng(123)
    out = []
    for k in range(12):
        n = int(rng.integers(50, 1200))
        kind = k % 4
        if kind == 0:    # overlapping blobs, negative and positive coordinates
            c = rng.uniform(-6, 6, (6, 2))
            X = np.concatenate([rng.normal(ci, 0.9, (n // 6, 2)) for ci in c])
        elif kind == 1:  # uniform near percolation
            X = rng.uniform(-8, 8, (n, 2))
        elif kind == 2:  # quantised coordinates: many exact-eps ties
            X = rng.integers(-20, 20, (n, 2)) * 0.25
        else:            # few tight blobs plus noise
            X = np.concatenate([rng.normal(rng.uniform(-3, 3, 2), 0.05, (n // 2, 2)),
                                rng.uniform(-4, 4, (n - n // 2, 2))])
        eps = float(rng.choice([0.25, 0.3, 0.5, 0.75]))
        ms = int(rng.choice([1, 3, 5, 10]))
        out.append((X.astype(np.float32), eps, ms))
    return out


@pytest.mark.parametrize("case", range(12))
def test_random_inputs_match_bruteforce(case):
    X, eps, ms = _random_inputs()[case]
    assert np.array_equal(oracle.dbscan2d(X, eps, ms), bruteforce_dbscan(X, eps, ms))


def test_reference_bruteforce_dense():
    """The reference's own dense brute-force case (tests/test_eps_boundary.py:174-191):
    0 disagreements after canonical relabelling."""
    rng = np.random.default_rng(7)
    centers = rng.uniform(-4.0, 4.0, size=(8, 2)).astype(np.float32)
    X = np.concatenate([rng.normal(loc=c, scale=1.0, size=(150, 2)).astype(np.float32)
                        for c in centers])
    rng.shuffle(X)
    ours = oracle.dbscan2d(X, 0.35, 5)
    assert np.array_equal(ours, bruteforce_dbscan(X, 0.35, 5))
    # the reference's brute force breaks border ties by original index; random
    # data has no exact ties, so the partitions agree
    assert np.array_equal(canonical(ours), canonical(bruteforce_dbscan(X, 0.35, 5)))


def test_special_cases():
    assert oracle.dbscan2d(np.zeros((0, 2), np.float32), 1.0, 5).shape == (0,)
    one = np.array([[1.0, 2.0]], np.float32)
    assert oracle.dbscan2d(one, 1.0, 1).tolist() == [0]
    assert oracle.dbscan2d(one, 1.0, 2).tolist() == [-1]
    rng = np.random.default_rng(0)
    X = rng.normal(0, 0.3, (30, 2)).astype(np.float32)
    assert (oracle.dbscan2d(X, 1.0, 10_000) == -1).all()
    # min_samples is narrowed to int32 (dbscan_cpu.cpp:53): 2^32 + 3 acts as 3,
    # 2^31 becomes negative and every point is noise (dbscan_grid.hpp:141)
    assert np.array_equal(oracle.dbscan2d(X, 1.0, 2**32 + 3), oracle.dbscan2d(X, 1.0, 3))
    assert (oracle.dbscan2d(X, 1.0, 2**31) == -1).all()
    # ms = 1 on isolated points: each point is its own cluster, numbered in
    # signed cell-key order (kx ascending)
    far = np.array([[5.0, 0.0], [-5.0, 0.0], [0.0, 0.0]], np.float32)
    assert oracle.dbscan2d(far, 1.0, 1).tolist() == [2, 0, 1]


def test_permutation_equivariance():
    X, eps, ms = W.blobs10k()
    base = oracle.dbscan2d(X, eps, ms)
    perm = np.random.default_rng(5).permutation(len(X))
    assert np.array_equal(oracle.dbscan2d(X[perm], eps, ms), base[perm])


def test_thread_count_independent():
    X, eps, ms = W.blobs10k()
    before = oracle.set_threads(0)
    try:
        oracle.set_threads(1)
        a = oracle.dbscan2d(X, eps, ms)
        oracle.set_threads(4)
        b = oracle.dbscan2d(X, eps, ms)
    finally:
        oracle.set_threads(before)
    assert np.array_equal(a, b)


def test_blobs10k_core_and_noise_match_sklearn():
    """Core and noise sets equal sklearn's except where sklearn's float64
    distance and the reference's float32 rule (or its own-cell rule) disagree
    on a pair at the eps boundary."""
    sk = pytest.importorskip("sklearn.cluster")
    X, eps, ms = W.blobs10k()
    dbg = oracle.dbscan2d_debug(X, eps, ms)
    core = np.zeros(len(X), bool)
    core[dbg["perm"]] = dbg["is_core"].astype(bool)
    model = sk.DBSCAN(eps=eps, min_samples=ms).fit(X)
    sk_core = np.zeros(len(X), bool)
    sk_core[model.core_sample_indices_] = True
    mismatch = np.flatnonzero(core != sk_core)
    if mismatch.size:
        # each disagreement must involve a pair whose float64 and float32
        # verdicts differ; count neighbours both ways for those points
        X64 = X.astype(np.float64)
        for i in mismatch:
            d64 = ((X64 - X64[i]) ** 2).sum(1)
            n64 = int((np.sqrt(d64) <= eps).sum())
            assert (n64 >= ms) == sk_core[i]
        assert mismatch.size <= 5
    noise = dbg["labels"] == -1
    sk_noise = model.labels_ == -1
    assert (noise != sk_noise).sum() <= mismatch.size * ms


@pytest.mark.skipif(not oracle.ref_available(), reason="reference not built (make -C oracle ref)")
@pytest.mark.parametrize("case", range(12))
def test_random_inputs_match_compiled_reference(case):
    X, eps, ms = _random_inputs()[case]
    assert np.array_equal(oracle.dbscan2d(X, eps, ms), oracle.ref_dbscan2d(X, eps, ms))


def test_planted_mistakes_are_caught():
    """Mutations of the rules (applied to the brute force, on the CPU) are
    visible to the fixtures the oracle is pinned with."""
    cases = constructed_cases()
    X, eps, ms = cases["border_tie_offset"]
    good = bruteforce_dbscan(X, eps, ms)
    # (d2, original index) tiebreak -- the reference CUDA rule -- picks the
    # other cluster for the tie point
    d2 = ((X - X[5]) ** 2).sum(1)
    core = good != -1
    core[5] = False
    cand = np.flatnonzero(core & (d2 <= np.float32(eps) ** 2))
    by_index = good[cand[np.lexsort((cand, d2[cand]))[0]]]
    assert by_index != good[5]
    # unsigned key order renumbers clusters on the negative-kx side
    X, eps, ms = cases["signed_kx_order"]
    lab = bruteforce_dbscan(X, eps, ms)
    kx, ky = W.cell_keys_np(X, eps)
    ukey = ((kx & 0xFFFFFFFF).astype(np.uint64) << np.uint64(32)) | (ky & 0xFFFFFFFF).astype(np.uint64)
    first = {}
    for c in np.unique(lab[lab >= 0]):
        first[c] = ukey[lab == c].min()
    unsigned_order = sorted(first, key=first.get)
    assert unsigned_order != sorted(first)


def test_fast_row_shuffle_is_generator_shuffle():
    """The workloads gather rows by a shuffled index instead of shuffling the
    2-D array in place; both consume the generator identically."""
    from tests.workloads import _shuffle_rows

    for n in (1, 2, 17, 5000):
        X = np.random.default_rng(n).normal(size=(n, 2)).astype(np.float32)
        a = X.copy()
        r1 = np.random.default_rng(99)
        r1.shuffle(a)
        r2 = np.random.default_rng(99)
        assert np.array_equal(_shuffle_rows(r2, X), a)
        assert r1.integers(1 << 30) == r2.integers(1 << 30)
