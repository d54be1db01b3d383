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


# ------------------------------------------------ command-line input / output

def test_rust_lines_documented_examples():
    # the examples of Rust's str::lines documentation
    assert oracle.rust_lines("foo\r\nbar\n\nbaz\n") == ["foo", "bar", "", "baz"]
    assert oracle.rust_lines("foo\nbar\n\r\nbaz") == ["foo", "bar", "", "baz"]
    assert oracle.rust_lines("foo\nbar\n\r\nbaz\r") == ["foo", "bar", "", "baz\r"]
    assert oracle.rust_lines("") == []
    assert oracle.rust_lines("\n") == [""]


@pytest.mark.parametrize("case", oracle.CHECK_CASES)
def test_load_tsv_recovers_check_py_ids(case):
    """check.py:17-36 writes '-', '', 'LNF', 'PLOT3' for missing and 'INF-v'
    or 'v' for allele v, and expects distances of A (check.py:39-44): the
    load must give A itself (README.md:81: '-', empty and non-numeric calls
    are missing, INF- is stripped), with LF and CRLF line ends."""
    n, L, seed, hashed = case
    A, cells = oracle.check_make(n, L, seed, hashed)
    text = "FILE\t" + "\t".join(f"L{j}" for j in range(L)) + "\n"
    text += "".join(f"S{i}\t" + "\t".join(row) + "\n" for i, row in enumerate(cells))
    for t in (text, text.replace("\n", "\r\n")):
        ids, calls = oracle.load_tsv(t)
        assert ids == [f"S{i}" for i in range(n)]
        assert np.array_equal(calls, A.astype(np.uint64))
    # -r keeps INF-v as a non-numeric call: exactly those cells become missing
    _, calls_r = oracle.load_tsv(text, skip_repl=True)
    inf = np.array([[c.startswith("INF-") for c in row] for row in cells])
    assert inf.any()
    assert np.array_equal(calls_r, np.where(inf, 0, A).astype(np.uint64))


def test_parse_allele_u64_range():
    assert oracle.parse_allele(str(2**64 - 1)) == 2**64 - 1
    assert oracle.parse_allele("007") == 7
    with pytest.raises(oracle.AlleleTooLarge):
        oracle.parse_allele(str(2**64))
    with pytest.raises(oracle.AlleleTooLarge):
        oracle.parse_allele("INF-" + "9" * 25)


def test_rows_padded_truncated_and_empty_lines():
    text = "id\tA\tB\tC\n\nx\t1\ny\t1\t2\t3\t4\t5\n\nz\n"
    ids, calls = oracle.load_tsv(text)
    assert ids == ["x", "y", "z"]
    assert calls.tolist() == [[1, 0, 0], [1, 2, 3], [0, 0, 0]]
    # a cell past the header width is still parsed: its overflow is an error
    with pytest.raises(oracle.AlleleTooLarge):
        oracle.load_tsv("id\tA\nx\t1\t" + "9" * 21 + "\n")


def test_completeness_filters_f64_and_order():
    from fractions import Fraction
    # (57 / 100) * 100 is 56.99999999999999 in binary64: a locus called in 57
    # of 100 samples fails -L 57, although 57% are called
    calls = np.zeros((100, 2), dtype=np.uint64)
    calls[:57, 0] = 3
    calls[:, 1] = 4
    assert (57 / 100) * 100 < 57
    _, c = oracle.completeness_filters([str(i) for i in range(100)], calls, locus_pct=57)
    assert c.shape == (100, 1)
    _, c = oracle.completeness_filters([str(i) for i in range(100)], calls, locus_pct=56.9)
    assert c.shape == (100, 2)
    # away from binary64 rounding the rule is the exact rational one
    rng = np.random.default_rng(5)
    for _ in range(200):
        n, L = rng.integers(1, 30, 2)
        X = (rng.random((n, L)) < rng.random()).astype(np.uint64)
        thr = float(rng.integers(0, 101)) + 0.5
        ids, Y = oracle.completeness_filters(list(range(n)), X, locus_pct=thr)
        keep = [k for k in range(L) if Fraction(int(X[:, k].sum()) * 100, int(n)) >= Fraction(thr)]
        assert np.array_equal(Y, X[:, keep])
    # the locus filter runs first, the sample filter then sees only kept loci
    X = np.array([[1, 0, 0], [1, 0, 0], [1, 5, 0]], dtype=np.uint64)
    ids, Y = oracle.completeness_filters(["a", "b", "c"], X, locus_pct=50, sample_pct=100)
    assert ids == ["a", "b", "c"] and Y.shape == (3, 1)
    ids, Y = oracle.completeness_filters(["a", "b", "c"], X, sample_pct=50)
    assert ids == ["c"] and Y.shape == (1, 3)
    # no samples: 0 / 0 is NaN, every locus is dropped
    ids, Y = oracle.completeness_filters([], np.zeros((0, 4), dtype=np.uint64), locus_pct=0)
    assert Y.shape == (0, 0)


def test_render_matrix_known_text():
    H = np.array([[0, 3, 9], [3, 0, 1], [9, 1, 0]])
    ids = ["a", "b", "c"]
    assert oracle.render_matrix(ids, H) == \
        "cgmlst-hamming-gpu\ta\tb\tc\na\t0\t3\t9\nb\t3\t0\t1\nc\t9\t1\t0\n"
    assert oracle.render_matrix(ids, H, mode="lower-tri", cap=4, osep="::", index_name="x") == \
        "x::a::b::c\na::0::0::0\nb::3::0::0\nc::4::1::0\n"
    assert oracle.render_matrix(ids, H, mode="upper-tri") == \
        "cgmlst-hamming-gpu\ta\tb\tc\na\t0\t3\t9\nb\t0\t0\t1\nc\t0\t0\t0\n"
    assert oracle.render_matrix([], np.zeros((0, 0))) == "cgmlst-hamming-gpu\n"


def test_rust_trim_is_white_space():
    """main.rs:99 trims with str::trim, i.e. char::is_whitespace, the Unicode
    White_Space property.  Python's str.isspace covers exactly White_Space
    plus the four information separators U+001C..U+001F (bidi class B/S), so
    the oracle's set must be isspace minus those four, over every code point."""
    py_space = {c for c in range(0x110000) if chr(c).isspace()}
    assert {ord(c) for c in oracle.RUST_WHITE_SPACE} == py_space - {0x1C, 0x1D, 0x1E, 0x1F}
    assert oracle.parse_allele("\x1c5") == 0 and oracle.parse_allele("5\x1f") == 0
    assert oracle.parse_allele("\x855\u3000") == 5 and oracle.parse_allele("\u200b5") == 0
    # a planted mistake (Python's own strip) fails the first check
    assert "\x1c5".strip() == "5"
