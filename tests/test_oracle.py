"""Pins of the CPU oracle against the reference's own known answers and
against routes that do not share its formula (CPU only).

- KAT: tensorcore/README.md:52 publishes the upper-triangle checksum
  85724a9d7d090f9f (algorithm tensorcore/tc_hamming.cu:241) of the synthetic
  30,000 x 5,000 set.  That set is the first 30,000 rows of
  python/gen_profiles.py's 50,000-row output (default_rng(1)); the oracle's
  brute force reproduces the checksum exactly.
- Edge vector: python/relu_hamming.py:56-57.
- Independent route: H = P P^T - O O^T - R with a materialised one-hot and
  integer matrix products, at several thresholds, on the check.py datasets.
- Hand-counted tiny cases, invariants (symmetry, zero diagonal,
  H <= min(called_i, called_j), permutation and relabelling invariance).
- Planted mistakes: a wrong-sign / dropped-term variant of the formula is
  shown to break the checks above (so the checks can fail).
"""
import numpy as np
import pytest

import oracle


def test_relu_edge_vector():
    a, b, want = oracle.relu_edge_vector()
    X = np.stack([a, b]).astype(np.uint64)
    R = oracle.dense_ranks(X)
    H = oracle.brute_full(R)
    assert H[0, 1] == want.sum() == H[1, 0]
    for k in range(len(a)):  # locus by locus
        Hk = oracle.brute_full(oracle.dense_ranks(X[:, k:k + 1]))
        assert Hk[0, 1] == want[k]


def test_hand_counted():
    X = np.array([[1, 2, 0, 3],
                  [1, 3, 2, 0],
                  [0, 0, 0, 0],
                  [2, 2, 2, 3]], dtype=np.uint16)
    H = oracle.brute_full(X)
    want = np.array([[0, 1, 0, 1], [1, 0, 0, 2], [0, 0, 0, 0], [1, 2, 0, 0]])
    # pair (0,1): loci 0 equal, 1 differ, 2 missing, 3 missing -> 1
    # pair (0,3): 1 vs 2 differ, 2 vs 2 equal, missing, 3 vs 3 equal -> 1
    # pair (1,3): 1 vs 2, 3 vs 2, 2 vs 2 equal, missing -> 2
    assert np.array_equal(H, want)


@pytest.mark.parametrize("case", oracle.CHECK_CASES)
def test_brute_matches_check_ref_and_one_hot_route(case):
    n, L, seed, hashed = case
    A, _ = oracle.check_make(n, L, seed, hashed)
    ref = oracle.check_ref(A)  # numpy restatement of check.py:39-44
    X = oracle.ranks_checkpy_style(A)
    H = oracle.brute_full(X).astype(np.int64)
    assert np.array_equal(H, ref)
    # relabelling (dense ranks from the u64 ids instead of first appearance)
    assert np.array_equal(oracle.brute_full(oracle.dense_ranks(A.astype(np.uint64))).astype(np.int64), ref)
    # independent route through the one-hot identity, at several thresholds
    sub = slice(0, 300)
    for tau in (1, 2, 7, 64, 10**6):
        assert np.array_equal(oracle.one_hot_split(X[sub], tau), ref[sub, sub])


def test_invariants():
    rng = np.random.default_rng(0)
    X = rng.integers(0, 6, size=(120, 50)).astype(np.uint16)
    H = oracle.brute_full(X).astype(np.int64)
    assert np.array_equal(H, H.T) and not np.diag(H).any()
    called = (X != 0).sum(1)
    assert (H <= np.minimum.outer(called, called)).all()
    perm = rng.permutation(120)
    assert np.array_equal(oracle.brute_full(X[perm]), H[np.ix_(perm, perm)])
    lperm = rng.permutation(50)
    assert np.array_equal(oracle.brute_full(X[:, lperm]), H)
    # no missing data: plain Hamming distance
    Y = rng.integers(1, 4, size=(60, 30)).astype(np.uint16)
    HY = oracle.brute_full(Y)
    assert np.array_equal(HY, (Y[:, None, :] != Y[None, :, :]).sum(-1))


def test_upper_block_and_checksum_consistent():
    rng = np.random.default_rng(1)
    X = rng.integers(0, 9, size=(333, 77)).astype(np.uint16)
    H = oracle.brute_full(X)
    cells = np.concatenate([H[i, i + 1:] for i in range(333)])
    assert np.array_equal(oracle.upper_block(X, 0, 333), cells)
    assert np.array_equal(oracle.upper_block(X, 100, 210), np.concatenate([H[i, i + 1:] for i in range(100, 210)]))
    h = 0
    for v in cells.tolist():
        h = (h * 1000003 + v) & 0xFFFFFFFFFFFFFFFF
    assert oracle.checksum_upper(X) == h == oracle.checksum_cells(cells)


def test_planted_mistakes_are_caught():
    """Variants with a dropped term or a wrong sign fail the pins."""
    A, _ = oracle.check_make(300, 97, 1, False)
    ref = oracle.check_ref(A)
    X = oracle.ranks_checkpy_style(A).astype(np.int64)
    P = X > 0
    dropped_missing = np.array([((X[i] != X)).sum(1) for i in range(300)])      # ignores [a!=0][b!=0]
    one_sided = np.array([((X[i] != X) & P[i]).sum(1) for i in range(300)])      # drops [b!=0]
    assert not np.array_equal(dropped_missing, ref)
    assert not np.array_equal(one_sided, ref)
    D, O, Pm, groups, _ = oracle.encode_reference(X[:200], 8)
    no_rare = Pm @ Pm.T - O @ O.T                                              # drops R
    np.fill_diagonal(no_rare, 0)
    assert not np.array_equal(no_rare, ref[:200, :200])


def test_gen_profiles_shape():
    P = oracle.gen_profiles(50, 40)
    assert P.shape == (50, 40) and P.min() >= 0 and P.max() < 5000


@pytest.mark.slow
def test_checksum_kat_synthetic_30k():
    """tensorcore/README.md:52: synthetic 30,000 x 5,000 -> 85724a9d7d090f9f
    (about a minute on 8 cores)."""
    A = oracle.gen_profiles(50000, 5000)[:30000]
    X = oracle.ranks_checkpy_style(A)
    assert oracle.checksum_upper(X) == 0x85724a9d7d090f9f
