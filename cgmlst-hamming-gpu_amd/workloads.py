"""Seeded synthetic allele-call workloads (numpy PCG64).

Shapes follow the benchmark configurations of this project:
  small_zipf   n=2,000 x L=1,142, per-locus alleles U{1..400}, Zipf(1.3)
               frequencies, u32 ids, 2% missing i.i.d.
  long_tail    n=63,005 x L=1,142, per-locus allele counts log-uniform in
               [5, 3000], Zipf(1.1) frequencies, missing clustered by sample
               (rate ~ Beta(0.5, 15) per sample)
  founders     n x L, 50 founder alleles per locus, each profile copies a
               founder per locus, 1% mutated to a fresh id, 1% missing

  hash_ids_fast
               per-locus multiplicative relabelling of ids to 32 / 64 bits

Each returns u64 allele ids, n x L row-major, 0 = missing.
"""
from __future__ import annotations

import numpy as np


def _zipf_locus(rng, n, A, s, missing_mask):
    w = np.arange(1, A + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    r = np.searchsorted(cdf, rng.random(n), side="right") + 1
    r = np.minimum(r, A)
    r[missing_mask] = 0
    return r


def small_zipf(n=2000, L=1142, seed=1, id_bits=32):
    rng = np.random.default_rng(seed)
    A_k = rng.integers(1, 401, size=L)
    miss = rng.random((n, L)) < 0.02
    out = np.zeros((n, L), dtype=np.uint64)
    hi = np.uint64(2**id_bits - 1)
    for k in range(L):
        ranks = _zipf_locus(rng, n, int(A_k[k]), 1.3, miss[:, k])
        # injective random ids per (locus, rank): sample without replacement
        table = np.unique(rng.integers(1, int(hi), size=int(A_k[k]) * 2 + 8, dtype=np.uint64))
        rng.shuffle(table)
        table = table[: int(A_k[k])]
        col = np.zeros(n, dtype=np.uint64)
        nz = ranks > 0
        col[nz] = table[ranks[nz] - 1]
        out[:, k] = col
    return out


def long_tail(n=63005, L=1142, seed=2, a_min=5, a_max=3000, s=1.1):
    rng = np.random.default_rng(seed)
    A_k = np.rint(np.exp(rng.uniform(np.log(a_min), np.log(a_max), size=L))).astype(np.int64)
    rate = rng.beta(0.5, 15.0, size=n)
    miss = rng.random((n, L)) < rate[:, None]
    out = np.zeros((n, L), dtype=np.uint64)
    for k in range(L):
        out[:, k] = _zipf_locus(rng, n, int(A_k[k]), s, miss[:, k])
    return out


def zipf_iid(n, L, seed, s=1.2, a_min=5, a_max=3000, miss=0.02):
    """Long-tail alleles (log-uniform allele counts per locus, Zipf(s)
    frequencies) with i.i.d. missing calls."""
    rng = np.random.default_rng(seed)
    A_k = np.rint(np.exp(rng.uniform(np.log(a_min), np.log(a_max), size=L))).astype(np.int64)
    out = np.zeros((n, L), dtype=np.uint64)
    for k in range(L):
        out[:, k] = _zipf_locus(rng, n, int(A_k[k]), s, rng.random(n) < miss)
    return out


def batch_shapes(count=64, seed=4, n_range=(500, 20000), L_range=(800, 3000)):
    """(n, L, seed) of the independent datasets of the batching workload:
    n and L uniform in their ranges."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    Ls = rng.integers(L_range[0], L_range[1] + 1, size=count)
    return [(int(ns[i]), int(Ls[i]), seed * 1000 + i) for i in range(count)]


def founders(n=100000, L=5000, seed=3, nf=50, mut=0.01, miss=0.01):
    """Dense-frequent founder model: every locus has nf founder alleles (ids
    1..nf); each profile copies one of them per locus, a fraction `mut` of the
    calls mutate to a fresh id never seen elsewhere, a fraction `miss` is
    missing.  Generated in row blocks (n x L u64)."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, L), dtype=np.uint64)
    next_id = nf + 1
    for i0 in range(0, n, 4096):
        i1 = min(n, i0 + 4096)
        blk = rng.integers(1, nf + 1, size=(i1 - i0, L), dtype=np.uint64)
        m = rng.random(blk.shape) < mut
        cnt = int(m.sum())
        blk[m] = np.arange(next_id, next_id + cnt, dtype=np.uint64)
        next_id += cnt
        blk[rng.random(blk.shape) < miss] = 0
        out[i0:i1] = blk
    return out


def hashed_ids(calls, bits=64, seed=7):
    """Replace every non-zero id by an injective pseudo-random 32- or 64-bit
    id (equality is all that matters)."""
    rng = np.random.default_rng(seed)
    calls = np.asarray(calls, dtype=np.uint64)
    uniq, inv = np.unique(calls, return_inverse=True)
    hi = 2**bits - 1
    while True:
        table = rng.integers(1, hi, size=uniq.size, dtype=np.uint64, endpoint=True)
        if np.unique(table).size == uniq.size:
            break
    table[uniq == 0] = 0
    return table[inv].reshape(calls.shape)


def serving_sets(n_db, L, n_queries, seed, fresh=0.03, a_min=3, a_max=600):
    """A database and a query set drawn from one long-tail population; a
    fraction `fresh` of the query calls are ids the database never saw, and
    all ids are arbitrary 64-bit values.  -> (db_calls, query_calls) u64."""
    calls = zipf_iid(n_db + n_queries, L, seed, s=1.2, a_min=a_min, a_max=a_max)
    rng = np.random.default_rng(seed)
    q = calls[n_db:]
    new = rng.random(q.shape) < fresh
    q[new] = (10**12 + np.arange(int(new.sum()))).astype(np.uint64)
    calls = hashed_ids(calls, bits=64, seed=seed)
    return calls[:n_db], calls[n_db:]


def hash_ids_fast(calls, bits, seed):
    """Relabel every id of locus k as id * M_k modulo 2^bits with a random odd
    M_k: a bijection on each locus's ids (so equality is kept), 0 stays 0
    (missing), and the rank order of the new ids is scrambled."""
    rng = np.random.default_rng(seed)
    L = calls.shape[1]
    m = (rng.integers(0, 2**62, size=L, dtype=np.uint64) * np.uint64(2) + np.uint64(1))
    out = calls.astype(np.uint64) * m[None, :]  # wraps modulo 2^64
    if bits == 32:
        out &= np.uint64(0xFFFFFFFF)
    return out


def batch_dataset(n, L, seed, d):
    """Dataset d of the 64-set batching workload (shape from batch_shapes):
    Zipf(1.2) long tail, 2% missing, ids relabelled to 32 bits in even
    datasets and 64 bits in odd ones."""
    return hash_ids_fast(zipf_iid(n, L, seed, s=1.2), 32 if d % 2 == 0 else 64, seed + 17)
