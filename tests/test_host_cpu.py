"""Host-side pieces on the CPU: the C ABI library loads and exports every
symbol the header declares, the encoder matches the numpy restatement of
main.rs:366-444, dense ranks match main.rs:135-165, the tau rule of
main.rs:169-198 behaves (capacity injected through CGH_FREE_BYTES), and the
CLI's usage / device errors keep the reference's exit codes.  No kernel is
launched here."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "cgmlst_hamming.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"if", "return", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and (n.startswith("tch_") or n.startswith("cgh_"))} - {"tch_cb"})


def test_library_exports_header(cgh):
    lib = cgh.load()
    declared = header_functions()
    assert set(declared) == set(cgh.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", cgh.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}$", out, re.M), name


def test_dense_ranks_match_oracle(cgh, workloads):
    calls = workloads.small_zipf(500, 120, seed=4, id_bits=32)
    assert np.array_equal(cgh.dense_ranks(calls), oracle.dense_ranks(calls))
    calls64 = workloads.hashed_ids(calls, bits=64)
    assert calls64.max() > 2**40
    X = cgh.dense_ranks(calls64)
    assert np.array_equal(X, oracle.dense_ranks(calls64))
    # ranks are ascending in the id within each locus
    k = 3
    ids = calls64[:, k]
    nz = ids != 0
    order = np.argsort(ids[nz], kind="stable")
    assert (np.diff(X[nz, k][order].astype(int)) >= 0).all()


def test_dense_ranks_limit(cgh):
    calls = np.arange(1, 70001, dtype=np.uint64).reshape(-1, 1)
    with pytest.raises(cgh.CghError, match=r"locus 0 has 70000 distinct alleles \(limit 65535\)"):
        cgh.dense_ranks(calls)


@pytest.mark.parametrize("tau", [1, 2, 3, 8, 64, 100000])
@pytest.mark.parametrize("k_align", [256, 1024])
def test_encoder_matches_restatement(cgh, tau, k_align):
    A, _ = oracle.check_make(300, 97, 1, False)
    X = oracle.dense_ranks(A.astype(np.uint64))
    enc = cgh.Encoded(X, min(tau, 301), k_align)
    e = enc.enc
    D, O, P, groups, partners = oracle.encode_reference(X, tau)
    assert e.D == D and e.n_pad % 256 == 0 and (e.kw1 * 32) % k_align == 0 and (e.kw2 * 32) % k_align == 0
    Ob = np.unpackbits(enc.O.view(np.uint8), axis=1, bitorder="little")
    Pb = np.unpackbits(enc.P.view(np.uint8), axis=1, bitorder="little")
    assert np.array_equal(Ob[:300, :D], O) and not Ob[:, D:].any() and not Ob[300:].any()
    assert np.array_equal(Pb[:300, :97], P) and not Pb[:, 97:].any() and not Pb[300:].any()
    gstart, gmem, gpairs = enc.groups()
    assert e.n_groups == len(groups)
    for g, (_, _, mem) in enumerate(groups):
        assert np.array_equal(gmem[gstart[g]:gstart[g + 1]], mem)
        c = len(mem)
        assert gpairs[g + 1] - gpairs[g] == c * (c - 1) // 2
    start, js = enc.rare_csr()
    assert e.n_rare == sum(len(p) for p in partners) == js.size
    rows = enc.row_index()
    for i in range(300):
        assert sorted(js[start[i]:start[i + 1]].tolist()) == sorted(partners[i])
        assert sorted(rows[i].tolist()) == sorted(partners[i])


def test_choose_tau_rule(cgh, monkeypatch):
    A, _ = oracle.check_make(700, 410, 2, True)
    X = oracle.dense_ranks(A.astype(np.uint64))
    lines = []
    tau = cgh.choose_tau(X, 2 * 512 * 768 * 2, 10**12, log=lines.append)
    assert tau in [2 ** k for k in range(1, 13)] and len(lines) == 12
    assert cgh.choose_tau(X, 2 * 512 * 768 * 2, 1000) == 0  # nothing fits in 1 kB
    monkeypatch.setenv("CGH_FREE_BYTES", "123456789")
    assert cgh.free_vram(0) == 123456789


def test_upper_layout_helpers(cgh):
    n, S = 1000, 192
    total = 0
    for s, r0 in enumerate(range(0, n, S)):
        assert cgh.stripe_offset(n, S, s) == total
        total += min(S, n - r0) * (n - r0)
    assert cgh.upper_elems(n, S) == (total + 7) // 8 * 8


def test_cli_usage_errors(cgh):
    cli = cgh.CLI_PATH
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "--input <INPUT>" in r.stderr
    r = subprocess.run([cli, "-i", "x", "-m", "diagonal"], capture_output=True, text=True)
    assert r.returncode == 2 and "possible values: full, lower-tri, upper-tri" in r.stderr
    r = subprocess.run([cli, "-i", "x", "--tau", "-3"], capture_output=True, text=True)
    assert r.returncode == 2
    r = subprocess.run([cli, "-V"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("cgmlst-hamming-gpu ")
    r = subprocess.run([cli, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "--stripe" in r.stdout


def test_cli_device_and_separator_errors(cgh):
    if cgh.device_count() > 0:
        pytest.skip("a GPU is visible: the device-missing path cannot be reached")
    cli = cgh.CLI_PATH
    r = subprocess.run([cli, "-i", "/nonexistent"], capture_output=True, text=True)
    # main.rs:283's text, kept verbatim by the drop-in
    assert r.returncode == 1 and r.stderr == "error: CUDA device 0 not available\n"
    r = subprocess.run([cli, "-i", "/nonexistent", "-d", "ab"], capture_output=True, text=True)
    assert r.returncode == 1 and "input separator must be a single character" in r.stderr


def _bits(words, width):
    """u32 rows -> 0/1 float32 matrix of the first `width` bits (LSB-first)."""
    b = np.unpackbits(np.ascontiguousarray(words).view(np.uint8), axis=1, bitorder="little")
    return b[:, :width].astype(np.float32)


@pytest.mark.parametrize("tau", [1, 4, 32, 10**6])
def test_query_codec_identity(cgh, workloads, tau):
    """Queries encoded against a database must give, through
    H = |P_q & P_j| - |O_q & O_j| - #ranges of q holding j, exactly the
    oracle's brute-force distances of the two sets ranked together; the
    database's own encoding must be cgh_encode's of its dense ranks."""
    n_db, L, nq = 1500, 150, 97
    db_calls, q_calls = workloads.serving_sets(n_db, L, nq, seed=31)
    X_all = oracle.dense_ranks(np.vstack([db_calls, q_calls]))
    want = oracle.brute_rows(X_all, n_db, n_db + nq)[:, :n_db].astype(np.int64)

    codec = cgh.QueryCodec(db_calls, tau, want_encoding=True)
    assert codec.tau == min(tau, n_db + 1)
    own = cgh.Encoded(cgh.dense_ranks(db_calls), codec.tau, 256)
    e, f = codec.enc, own.enc
    assert (e.kw1, e.kw2, e.D) == (f.kw1, f.kw2, f.D) == (codec.kw1, codec.kw2, codec.D)
    assert np.array_equal(own.O, np.ctypeslib.as_array(e.O, shape=(e.n_pad, e.kw1)))
    assert np.array_equal(own.P, np.ctypeslib.as_array(e.P, shape=(e.n_pad, e.kw2)))

    Oq, Pq, rs, re = codec.encode(q_calls)
    assert rs[0] == 0 and rs[-1] == re.size and np.all(np.diff(rs.astype(np.int64)) >= 0)
    mem = codec.members()
    b, end = (re & 0xFFFFFFFF).astype(np.int64), (re >> np.uint64(32)).astype(np.int64)
    assert np.all(b <= end) and np.all(end <= mem.size)
    R = np.zeros((nq, n_db), dtype=np.int64)
    for r in range(nq):
        for k in range(int(rs[r]), int(rs[r + 1])):
            R[r, mem[b[k]:end[k]]] += 1
    Odb, Pdb = own.O[:n_db], own.P[:n_db]
    H = (_bits(Pq, L) @ _bits(Pdb, L).T).astype(np.int64) - (
        _bits(Oq, codec.D) @ _bits(Odb, codec.D).T).astype(np.int64) - R
    assert np.array_equal(H, want)
    # every list member carries the allele: lists partition the database's
    # non-frequent calls (count < tau, singletons included)
    assert mem.size == int((X_all[:n_db] > 0).sum()) - int(_bits(Odb, codec.D).sum())
    codec.close()


def test_query_codec_rows_match_stacked(cgh, workloads):
    """cgh_codec_encode_rows over queries held in separate buffers (the
    server's batch of requests) gives exactly cgh_codec_encode's output for
    the same queries stacked in one array: bits, row starts and entries."""
    db_calls, q_calls = workloads.serving_sets(800, 120, 61, seed=5)
    codec = cgh.QueryCodec(db_calls, 16)
    want = codec.encode(q_calls)
    parts = [q_calls[:1], q_calls[1:20], q_calls[20:21], q_calls[21:]]
    got = codec.encode_rows(parts)
    for w, g in zip(want, got):
        assert np.array_equal(w, g)
    codec.close()


def test_query_codec_rejects_short_buffer(cgh, workloads):
    db_calls, q_calls = workloads.serving_sets(200, 40, 5, seed=3)
    codec = cgh.QueryCodec(db_calls, 8)
    lib = cgh.load()
    q = np.ascontiguousarray(q_calls)
    O = np.zeros((5, codec.kw1), np.uint32)
    P = np.zeros((5, codec.kw2), np.uint32)
    rs = np.zeros(6, np.uint32)
    re = np.zeros(5 * 40 - 1, np.uint64)
    err = cgh._err()
    assert lib.cgh_codec_encode(codec.handle, cgh._ptr(q), 5, cgh._ptr(O), cgh._ptr(P), cgh._ptr(rs), cgh._ptr(re),
                                re.size, err, 512) == -1
    assert b"rare entries" in err.value


def test_bench_one_gpu_only():
    """bench.py measures one GPU: any other --gpus exits 2 with a message,
    before touching a device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("tau", [2, 8, 64, 10**6])
def test_member_range_count(cgh, tau):
    """The tau model counts one member range per member of a rare group but
    the group's last (sum of c - 1 over the allele counts 2 <= c < tau that
    cgh_choose_tau reads): that is the size of the encoder's row index."""
    A, _ = oracle.check_make(700, 410, 2, True)
    X = oracle.dense_ranks(A.astype(np.uint64))
    enc = cgh.Encoded(X, tau, 256)
    cnt = cgh.allele_counts(X).astype(np.int64)
    c = cnt[(cnt >= 2) & (cnt < tau)]
    assert int(enc.enc.n_rentries) == int((c - 1).sum())
