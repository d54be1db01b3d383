"""CPU oracle for the exact cgMLST allelic Hamming distance path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.

Contents and the reference lines each restates:
  check_make, check_ref, check_cases, check_options
      tensorcore/rs/check.py:17-36 (seeded awkward TSV generator, numpy PCG64),
      39-44 (numpy brute force), 58 (cases), 69-78 (option matrix)
  gen_profiles            python/gen_profiles.py:1-12 (50-founder generator)
  ranks_checkpy_style     tensorcore/to_ranks.py:3-11 semantics (only
                          equality matters; first-appearance order)
  brute_rows / checksum_upper / dense_ranks
                          ctypes over oracle.c (see its header for citations)
  one_hot_split           H = P P^T - O O^T - R with an explicit one-hot at a
                          threshold tau (README.md:20-60), an independent
                          route used to pin the brute force.
  encode_reference        main.rs:366-444 (column map, O/P bits LSB-first,
                          rare groups, CSR) in numpy, the golden of the
                          encoder.
  relu_edge_vector        python/relu_hamming.py:56-57 known answer.
  rust_lines, parse_allele, load_tsv, completeness_filters, render_matrix
                          main.rs:95-113 (cell semantics), 286-311 (load:
                          Rust str::lines, header width, empty lines dropped,
                          rows padded / truncated), 313-347 (-L then -S in
                          f64), 225-265 (text of the matrix), the expected
                          bytes of the CLI edge tests.

Pins: see tests/test_oracle.py (checksum KAT 85724a9d7d090f9f of
tensorcore/README.md:52, the edge vector, hand counts, invariants, and the
one-hot identity).  No function here is "parity unpinned".
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def use_library(path):
    """Load the oracle from another build of oracle.c (the CPU baseline uses
    one compiled for the host it runs on); later calls go to it."""
    global _lib
    _lib = None
    return lib(path)


def lib(path=None):
    global _lib
    if _lib is None:
        path = path or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: make -C oracle")
        L = C.CDLL(path)
        P, I, S = C.c_void_p, C.c_int, C.c_size_t
        L.oracle_hamming_rows.argtypes = [P, I, I, I, I, P]
        L.oracle_hamming_rows.restype = None
        L.oracle_upper_block.argtypes = [P, I, I, I, I, P]
        L.oracle_upper_block.restype = None
        L.oracle_upper_cells.argtypes = [I, I, I]
        L.oracle_upper_cells.restype = S
        L.oracle_checksum_upper.argtypes = [P, I, I]
        L.oracle_checksum_upper.restype = C.c_uint64
        L.oracle_checksum_cells.argtypes = [P, S, C.c_uint64]
        L.oracle_checksum_cells.restype = C.c_uint64
        L.oracle_dense_ranks.argtypes = [P, I, I, P]
        L.oracle_dense_ranks.restype = I
        L.oracle_num_threads.argtypes = []
        L.oracle_num_threads.restype = I
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- check.py

def check_make(n, L, seed, hashed):
    """check.py:17-36, call for call (numpy Generator(PCG64(seed)))."""
    rng = np.random.default_rng(seed)
    founders = rng.integers(1, 30, size=(8, L))
    A = founders[rng.integers(0, 8, n)]
    m = rng.random((n, L)) < 0.15
    A[m] = rng.integers(1, 2000, m.sum())
    A[rng.random((n, L)) < 0.05] = 0
    if hashed:
        A = np.where(A > 0, (A * 2654435761) % 4294967291 + 1, 0)
    cells = []
    for r in A:
        row = []
        for v in r:
            u = rng.random()
            if v == 0:
                row.append("-" if u < 0.4 else "" if u < 0.6 else "LNF" if u < 0.8 else "PLOT3")
            else:
                row.append(f"INF-{v}" if u < 0.1 else str(v))
        cells.append(row)
    return A, cells


def check_ref(A):
    """check.py:39-44: H[i] = ((A[i] != A) & P[i] & P).sum(1)."""
    P = A > 0
    H = np.zeros((len(A), len(A)), dtype=np.int64)
    for i in range(len(A)):
        H[i] = ((A[i] != A) & P[i] & P).sum(1)
    return H


CHECK_CASES = [(300, 97, 1, False), (700, 410, 2, True), (1500, 64, 3, False)]  # check.py:58


def check_options(H):
    """check.py:69-78: (CLI args, expected matrix)."""
    iu = np.triu(np.ones_like(H, dtype=bool))
    return [
        ([], H),
        (["--tau", "2"], H),
        (["--tau", "100000"], H),
        (["--stripe", "64"], H),
        (["-m", "lower-tri"], np.where(iu.T, H, 0)),
        (["-m", "upper-tri"], np.where(iu, H, 0)),
        (["-X", "7"], np.minimum(H, 7)),
        (["-D", ","], H),
    ]


def write_check_tsv(path, cells, L):
    """check.py:62-66 layout: header FILE\\tL0.., rows S{i}\\t..."""
    with open(path, "w") as f:
        f.write("FILE\t" + "\t".join(f"L{j}" for j in range(L)) + "\n")
        for i, row in enumerate(cells):
            f.write(f"S{i}\t" + "\t".join(row) + "\n")


def parse_cli_output(text, sep="\t"):
    """check.py:49-53: rows after the header, fields after the id, as ints."""
    lines = text.rstrip("\n").split("\n")
    return np.array([[int(x) for x in l.split(sep)[1:]] for l in lines[1:]])


# ---------------------------------------------------------- gen_profiles.py

def gen_profiles(n, L, seed=1):
    """python/gen_profiles.py:3-8: 50 founders from 1..199, 3% mutations to
    1..4999, 2% missing (0)."""
    rng = np.random.default_rng(seed)
    F = rng.integers(1, 200, size=(50, L))
    P = F[rng.integers(0, 50, n)]
    m = rng.random((n, L)) < 0.03
    P[m] = rng.integers(1, 5000, m.sum())
    P[rng.random((n, L)) < 0.02] = 0
    return P


def ranks_checkpy_style(A):
    """Per-locus ranks in first-appearance order (to_ranks.py:5-8 uses
    pd.factorize); only equality of ranks matters for H."""
    A = np.asarray(A)
    X = np.zeros(A.shape, dtype=np.uint16)
    for k in range(A.shape[1]):
        col = A[:, k]
        nz = col != 0
        _, first, inv = np.unique(col[nz], return_index=True, return_inverse=True)
        order = np.argsort(np.argsort(first))
        X[nz, k] = (order[inv] + 1).astype(np.uint16)
    return X


# ----------------------------------------------------------- C brute force

def brute_full(X):
    """Full n x n H from dense ranks (or any u16 ids), oracle.c."""
    X = np.ascontiguousarray(X, dtype=np.uint16)
    n, L = X.shape
    H = np.zeros((n, n), dtype=np.uint16)
    lib().oracle_hamming_rows(_p(X), n, L, 0, n, _p(H))
    return H


def brute_rows(X, i0, i1):
    X = np.ascontiguousarray(X, dtype=np.uint16)
    n, L = X.shape
    H = np.zeros((i1 - i0, n), dtype=np.uint16)
    lib().oracle_hamming_rows(_p(X), n, L, i0, i1, _p(H))
    return H


def upper_block(X, i0, i1):
    """Strict upper part of rows i0..i1-1 concatenated (row i: j = i+1..n-1)."""
    X = np.ascontiguousarray(X, dtype=np.uint16)
    n, L = X.shape
    cells = lib().oracle_upper_cells(n, i0, i1)
    out = np.zeros(max(cells, 1), dtype=np.uint16)
    lib().oracle_upper_block(_p(X), n, L, i0, i1, _p(out))
    return out[:cells]


def checksum_upper(X):
    """tc_hamming.cu:241 checksum of the strict upper triangle."""
    X = np.ascontiguousarray(X, dtype=np.uint16)
    n, L = X.shape
    return int(lib().oracle_checksum_upper(_p(X), n, L))


def checksum_cells(cells, h=0):
    cells = np.ascontiguousarray(cells, dtype=np.uint16)
    return int(lib().oracle_checksum_cells(_p(cells), cells.size, h))


def dense_ranks(calls):
    """main.rs:135-165 in C; raises like the reference dies."""
    calls = np.ascontiguousarray(calls, dtype=np.uint64)
    n, L = calls.shape
    X = np.zeros((n, L), dtype=np.uint16)
    r = lib().oracle_dense_ranks(_p(calls), n, L, _p(X))
    if r < 0:
        raise ValueError("locus has more than 65535 distinct alleles")
    return X


def num_threads():
    return lib().oracle_num_threads()


# ------------------------------------------------------ independent routes

def locus_counts(X):
    return [np.bincount(X[:, k], minlength=int(X[:, k].max()) + 1) for k in range(X.shape[1])]


def encode_reference(X, tau):
    """main.rs:366-444 in numpy: (D, O bits as 0/1 matrix n x D, P 0/1 n x L,
    rare groups [(locus, rank, members)], CSR start/js as sorted lists)."""
    X = np.asarray(X, dtype=np.int64)
    n, L = X.shape
    cnt = locus_counts(X)
    cols, base = [], [0]
    for k in range(L):
        c = np.full(cnt[k].size, -1, dtype=np.int64)
        freq = [a for a in range(1, cnt[k].size) if cnt[k][a] >= tau]
        c[freq] = np.arange(len(freq))
        cols.append(c)
        base.append(base[-1] + len(freq))
    D = base[-1]
    O = np.zeros((n, D), dtype=np.int64)
    P = (X != 0).astype(np.int64)
    groups = []
    for k in range(L):
        for a in range(1, cnt[k].size):
            if cols[k][a] >= 0:
                O[X[:, k] == a, base[k] + cols[k][a]] = 1
            elif cnt[k][a] >= 2:
                groups.append((k, a, np.nonzero(X[:, k] == a)[0]))
    partners = [[] for _ in range(n)]
    for _, _, mem in groups:
        for ia in range(len(mem)):
            partners[mem[ia]].extend(mem[ia + 1:].tolist())
    return D, O, P, groups, partners


def one_hot_split(X, tau):
    """H = P P^T - O O^T - R via integer matrix products (README.md:20-33)."""
    D, O, P, groups, _ = encode_reference(X, tau)
    C_ = P @ P.T
    E = O @ O.T
    n = X.shape[0]
    R = np.zeros((n, n), dtype=np.int64)
    for _, _, mem in groups:
        R[np.ix_(mem, mem)] += 1
    H = C_ - E - R
    np.fill_diagonal(H, 0)
    return H


def relu_edge_vector():
    """python/relu_hamming.py:56-57: per-locus mismatch bits of a vs b."""
    a = np.array([5, 5, 6, 0, 7, 0, 10**9])
    b = np.array([5, 6, 5, 3, 0, 0, 1])
    return a, b, np.array([0, 1, 1, 0, 0, 0, 1])


# ---------------------------------------------------------------------------
# Input semantics of the command line (main.rs:95-113, 286-347)
# ---------------------------------------------------------------------------

class AlleleTooLarge(ValueError):
    pass


def rust_lines(text):
    """str::lines as main.rs:287 uses it: split at LF, drop the CR of a CRLF,
    no final empty line after a trailing LF; a bare CR ending the text is
    kept (documented behaviour of str::lines since Rust 1.77)."""
    if text == "":
        return []
    parts = text.split("\n")
    ended = text.endswith("\n")
    if ended:
        parts = parts[:-1]
    last = len(parts) - 1
    return [p[:-1] if p.endswith("\r") and (i < last or ended) else p
            for i, p in enumerate(parts)]


# Rust's str::trim removes char::is_whitespace from both ends: the Unicode
# White_Space property (Rust core docs), which Python's str.strip does not
# follow (it also strips U+001C..U+001F).
RUST_WHITE_SPACE = ("\t\n\x0b\x0c\r \x85\xa0\u1680" + "".join(chr(c) for c in range(0x2000, 0x200B)) +
                    "\u2028\u2029\u202f\u205f\u3000")


def rust_trim(s):
    """str::trim (main.rs:99): strip White_Space characters from both ends."""
    return s.strip(RUST_WHITE_SPACE)


def parse_allele(cell, skip_repl=False, missing="-"):
    """main.rs:95-113: trim, strip one leading INF- (unless skip_repl), the
    missing value or empty -> 0, all ASCII digits -> u64 (overflow raises),
    anything else -> 0."""
    s = rust_trim(cell)
    if not skip_repl and s.startswith("INF-"):
        s = s[4:]
    if s == missing or s == "":
        return 0
    if all("0" <= ch <= "9" for ch in s):
        v = int(s)
        if v >= 1 << 64:
            raise AlleleTooLarge(s)
        return v
    return 0


def load_tsv(text, sep="\t", skip_repl=False, missing="-"):
    """main.rs:286-311 -> (ids, calls n x L u64).  L = header fields - 1;
    empty lines are dropped; every cell of a row is parsed, then the row is
    padded with 0 or cut to L."""
    lines = rust_lines(text)
    if not lines:
        raise ValueError("empty input")
    L = len(lines[0].split(sep)) - 1
    rows = [ln for ln in lines[1:] if ln != ""]
    ids, calls = [], np.zeros((len(rows), L), dtype=np.uint64)
    for i, ln in enumerate(rows):
        f = ln.split(sep)
        ids.append(f[0])
        vals = [parse_allele(c, skip_repl, missing) for c in f[1:]]
        k = min(L, len(vals))
        calls[i, :k] = np.array(vals[:k], dtype=np.uint64)
    return ids, calls


def completeness_filters(ids, calls, locus_pct=None, sample_pct=None):
    """main.rs:313-347: keep loci with (called / n) * 100 >= thr, then samples
    with (called / L) * 100 >= thr, in f64 and in that order."""
    n, L = calls.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        if locus_pct is not None:
            keep = [k for k in range(L)
                    if (np.float64(np.count_nonzero(calls[:, k])) / np.float64(n)) * 100.0 >= locus_pct]
            calls = calls[:, keep]
            L = calls.shape[1]
        if sample_pct is not None:
            keep = [i for i in range(n)
                    if (np.float64(np.count_nonzero(calls[i])) / np.float64(L)) * 100.0 >= sample_pct]
            calls = calls[keep]
            ids = [ids[i] for i in keep]
    return ids, calls


def render_matrix(ids, H, mode="full", cap=0, osep="\t", index_name="cgmlst-hamming-gpu"):
    """main.rs:225-265: header index_name + (osep id)*, then per row the id and
    every cell; cells outside the layout print 0, the cap applies when
    cap > 0 and v > cap; LF line ends."""
    n = len(ids)
    out = [index_name + "".join(osep + x for x in ids) + "\n"]
    for i in range(n):
        row = [ids[i]]
        for j in range(n):
            keep = mode == "full" or (mode == "lower-tri" and j <= i) or (mode == "upper-tri" and j >= i)
            v = int(H[i, j]) if keep else 0
            if cap > 0 and v > cap:
                v = cap
            row.append(str(v))
        out.append(osep.join(row) + "\n")
    return "".join(out)


def cli_expected(text, sep="\t", skip_repl=False, missing="-", locus_pct=None, sample_pct=None, **fmt):
    """Expected stdout of the command line for an input text, or raises the
    error main.rs dies with (AlleleTooLarge, ValueError)."""
    ids, calls = load_tsv(text, sep, skip_repl, missing)
    ids, calls = completeness_filters(ids, calls, locus_pct, sample_pct)
    if calls.shape[1] > 65535:
        raise ValueError("more than 65535 loci (u16 distances)")
    n, L = calls.shape
    H = brute_full(dense_ranks(calls)) if n and L else np.zeros((n, n), dtype=np.int64)
    return render_matrix(ids, H, **fmt)
