"""GPU parity: the HIP path against the oracle, bit-exact.

Every test here runs the gfx950 kernels through the C ABI (or the CLI that
calls it) and compares every cell with the oracle's brute force
Σ_k [a≠0][b≠0][a≠b] (tensorcore/rs/check.py:39-44).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


KERNELS = [1, 2, 3]  # CGH_KERNEL_MFMA_I8, CGH_KERNEL_MFMA_FP4, CGH_KERNEL_VALU_POPC


def gpu_upper(cgh, torch, X, tau, S, k_align=256, kernel=0, cap=0):
    """Encode X at tau, upload, run cgh_dataset_upper, return the square H."""
    enc = cgh.Encoded(X, tau, k_align)
    ds = cgh.Dataset(enc)
    if cap:
        ds.set_cap(cap)
    n = X.shape[0]
    out = torch.full((cgh.upper_elems(n, S),), 0x5a5a, dtype=torch.int16, device="cuda")
    st = torch.cuda.current_stream()
    ds.upper(S, out.data_ptr(), st.cuda_stream, kernel)
    torch.cuda.synchronize()
    flat = out.cpu().numpy().view(np.uint16)
    ds.close()
    return cgh.upper_to_square(flat, n, S), flat, enc


@pytest.mark.parametrize("case", oracle.CHECK_CASES)
def test_check_matrix_cli(cgh, gpu, case):
    """The reference's own exactness matrix (check.py:56-83): 3 datasets x 8
    option sets through the drop-in CLI, every cell."""
    n, L, seed, hashed = case
    A, cells = oracle.check_make(n, L, seed, hashed)
    H = oracle.check_ref(A)
    assert np.array_equal(H, oracle.brute_full(oracle.ranks_checkpy_style(A)).astype(np.int64))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, f"c{seed}.tsv")
        oracle.write_check_tsv(path, cells, L)
        for args, want in oracle.check_options(H):
            print(f"case n={n} L={L} seed={seed} args={args}", flush=True)
            out = subprocess.run([cgh.CLI_PATH, "-s", "-i", path, *args], capture_output=True, text=True,
                                 timeout=120)
            assert out.returncode == 0, out.stderr
            got = oracle.parse_cli_output(out.stdout, "," if "-D" in args else "\t")
            assert np.array_equal(got, want), (case, args)


def test_cli_gzip_and_file_outputs(cgh, gpu):
    """Output written while the stripes arrive: to a file, gzip (parallel
    members) by -z or by a .gz name, multi-stripe, every layout; the
    decompressed text must match the oracle cell for cell."""
    import gzip
    n, L, seed, hashed = oracle.CHECK_CASES[-1]
    A, cells = oracle.check_make(n, L, seed, hashed)
    H = oracle.check_ref(A)
    iu = np.triu(np.ones_like(H, dtype=bool))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "in.tsv")
        oracle.write_check_tsv(path, cells, L)
        runs = [
            (["-z", "-o", os.path.join(d, "a")], True, H),
            (["-o", os.path.join(d, "b.gz"), "--stripe", "64", "-m", "upper-tri"], True, np.where(iu, H, 0)),
            (["-o", os.path.join(d, "c.tsv"), "--stripe", "8", "-m", "lower-tri", "-X", "9"], False,
             np.minimum(np.where(iu.T, H, 0), 9)),
        ]
        for args, zipped, want in runs:
            print(f"cli outputs args={args[2:]}", flush=True)
            out = subprocess.run([cgh.CLI_PATH, "-s", "-i", path, *args], capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stderr
            target = args[args.index("-o") + 1]
            raw = open(target, "rb").read()
            assert (raw[:2] == b"\x1f\x8b") == zipped
            text = (gzip.decompress(raw) if zipped else raw).decode()
            assert np.array_equal(oracle.parse_cli_output(text), want), args


@pytest.mark.parametrize("kernel", [2, 3])
@pytest.mark.parametrize("tau", [1, 64, 4096])
@pytest.mark.parametrize("cap", [1, 9, 400])
def test_dataset_cap_on_device(cgh, gpu, workloads, tau, cap, kernel):
    """-X on the device: min(H, cap) for every cell, whether the row is
    finished by the product epilogue (no rare partners) or by the rare pass."""
    X = cgh.dense_ranks(workloads.small_zipf(1200, 500, seed=3))
    H_ref = oracle.brute_full(X)
    print(f"dataset cap={cap} tau={tau} kernel={kernel}", flush=True)
    H, _, _ = gpu_upper(cgh, gpu, X, tau, 256, kernel=kernel, cap=cap)
    assert np.array_equal(H, np.minimum(H_ref, cap))


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("tau", [1, 2, 8, 64, 4096])
@pytest.mark.parametrize("S", [64, 512, 2048])
def test_dataset_small_zipf(cgh, gpu, workloads, tau, S, kernel):
    """2,000 x 1,142 long-tail profiles, every cell, across tau and stripes."""
    calls = workloads.small_zipf(2000, 1142, seed=1)
    X = cgh.dense_ranks(calls)
    H_ref = oracle.brute_full(X)
    print(f"dataset small_zipf tau={tau} S={S} kernel={kernel}", flush=True)
    H, flat, enc = gpu_upper(cgh, gpu, X, tau, S, kernel=kernel)
    assert np.array_equal(H, H_ref)


def test_dataset_lower_cells_zero(cgh, gpu, workloads):
    """Cells with j <= i inside each stripe are written as 0."""
    calls = workloads.small_zipf(700, 300, seed=3)
    X = cgh.dense_ranks(calls)
    n, S = X.shape[0], 256
    _, flat, _ = gpu_upper(cgh, gpu, X, 16, S)
    for s, r0 in enumerate(range(0, n, S)):
        rows, w = min(S, n - r0), n - r0
        off = cgh.stripe_offset(n, S, s)
        blk = flat[off:off + rows * w].reshape(rows, w)
        assert not np.tril(blk[:, :rows]).any()


@pytest.mark.parametrize("kernel", KERNELS)
def test_dataset_tile256(cgh, gpu, kernel):
    """n >= 8192 takes the 256 x 256 tile: random ids with missing data."""
    rng = np.random.default_rng(5)
    n, L = 9000, 200
    X = rng.integers(0, 40, size=(n, L)).astype(np.uint16)
    X[rng.random((n, L)) < 0.1] = 0
    print(f"dataset tile256 kernel={kernel}", flush=True)
    H, _, _ = gpu_upper(cgh, gpu, X, 32, 4096, kernel=kernel)
    assert np.array_equal(H, oracle.brute_full(X))


@pytest.mark.parametrize("shape", [(1500, 700, 64), (9000, 200, 4096)])
def test_autotune_then_exact(cgh, gpu, workloads, shape):
    """The autotuner picks one of the timed candidates and the dataset then
    runs with it (CGH_KERNEL_AUTO), every cell exact."""
    n, L, S = shape
    X = cgh.dense_ranks(workloads.small_zipf(n, L, seed=11))
    ds = cgh.Dataset(cgh.Encoded(X, 32, 256))
    print(f"autotune n={n} S={S}", flush=True)
    k, t, ms = ds.autotune(S)  # raises if any of the 12 candidates disagrees with the first
    assert k in (1, 2) and t in (128, 256) and ms > 0
    assert ds.tile(0) == t
    tk, tt, stages = ds.tuning()
    assert (tk, tt) == (k, t) and stages in (2, 3, 4)
    out = gpu.zeros(cgh.upper_elems(n, S), dtype=gpu.int16, device="cuda")
    ds.upper(S, out.data_ptr(), gpu.cuda.current_stream().cuda_stream, 0)
    gpu.cuda.synchronize()
    H = cgh.upper_to_square(out.cpu().numpy().view(np.uint16), n, S)
    assert np.array_equal(H, oracle.brute_full(X))
    ds.close()


@pytest.mark.parametrize("force_host", [False, True])
@pytest.mark.parametrize("shape", [(700, 410, 2, 64, 64), (1500, 64, 3, 2, 512), (1500, 64, 3, 4096, 8)])
def test_tch_run(cgh, gpu, monkeypatch, force_host, shape):
    """The reference-compatible tch_run entry point on check.py's datasets,
    two shapes x stripe heights (S = 8 makes a stripe narrower than a
    tile), rare pairs on the GPU and (free memory forced to 1 MiB) on the
    host, each row's partners in a shuffled order."""
    n, L, seed, tau, S = shape
    A, _ = oracle.check_make(n, L, seed, True)
    X = oracle.ranks_checkpy_style(A)
    if force_host:
        monkeypatch.setenv("CGH_FREE_BYTES", str(1 << 20))
    enc = cgh.Encoded(X, tau, 1024)
    print(f"tch_run n={n} tau={tau} S={S} force_host={force_host}", flush=True)
    H, secs, rare_gpu, seen = cgh.tch_run(enc, S=S, shuffle_seed=seed)
    assert rare_gpu == (not force_host or enc.enc.n_rare == 0)
    assert [s[0] for s in seen] == list(range(0, n, S))
    assert [s[2] for s in seen] == [n - r0 for r0 in range(0, n, S)]
    assert secs > 0
    assert np.array_equal(H, oracle.check_ref(A).astype(np.uint16))


def test_tch_run_rows_wider_than_one_chunk(cgh, gpu):
    """tch_run with n = 70,000: rows with more than 65,536 cells right of
    the diagonal are staged in two column chunks, and a CSR row's partners
    (here shuffled) are walked whole in each chunk with the members outside
    it masked; rows on both sides of the split against the oracle."""
    rng = np.random.default_rng(23)
    n, L = 70000, 4
    X = rng.integers(1, 2000, size=(n, L)).astype(np.uint16)  # ~35 copies per allele: all rare at tau 64
    X[rng.random((n, L)) < 0.03] = 0
    enc = cgh.Encoded(X, 64, 1024)
    assert enc.enc.n_rare > 10**6
    rows = [0, 1, 4462, 4463, 4464, 30000, 69998]
    print("tch_run n=70000 S=4096", flush=True)
    got, _, rare_gpu, _ = cgh.tch_run(enc, S=4096, keep_rows=rows, shuffle_seed=5)
    assert rare_gpu
    for i in rows:
        assert np.array_equal(got[i], oracle.brute_rows(X, i, i + 1)[0, i + 1:]), i


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("S,tau", [(64, 8), (520, 64), (4096, 1024), (8, 2)])
def test_dataset_stream(cgh, gpu, workloads, S, tau, graph):
    """cgh_dataset_stream (the CLI's path) and its HIP-graph form: stripes in
    order, every cell."""
    calls = workloads.small_zipf(1500, 700, seed=9)
    X = cgh.dense_ranks(calls)
    enc = cgh.Encoded(X, tau, 256)
    ds = cgh.Dataset(enc)
    print(f"dataset_stream S={S} tau={tau} graph={graph}", flush=True)
    H, secs, seen = ds.stream(S, graph)
    ds.close()
    assert [s[0] for s in seen] == list(range(0, 1500, S))
    assert np.array_equal(H, oracle.brute_full(X))


@pytest.mark.parametrize("kernel", KERNELS)
def test_batch_matches_oracle_and_graph_replays(cgh, gpu, workloads, kernel):
    """Many uneven datasets in one persistent launch: every dataset's layout
    equals the oracle; HIP-graph replays (with eager work in between) repeat
    the eager result exactly."""
    shapes = [(1, 50, 1), (2, 40, 2), (300, 97, 3), (700, 410, 4), (1500, 64, 5), (5000, 300, 6), (2600, 900, 7)]
    S = 512
    sets, Xs = [], []
    for n, L, seed in shapes:
        calls = workloads.zipf_iid(n, L, seed)
        if seed % 2:
            calls = workloads.hashed_ids(calls, bits=64 if seed % 3 else 32, seed=seed)
        X = cgh.dense_ranks(calls)
        Xs.append(X)
        tau = [1, 2, 8, 64, 4096, 32, 16][seed - 1]
        sets.append(cgh.Dataset(cgh.Encoded(X, tau, 256)))
    b = cgh.Batch(sets, S)
    assert [b.tile(d) for d in range(len(shapes))] == [128, 128, 128, 128, 128, 256, 128]
    out = gpu.full((b.out_elems(),), 0x3c3c, dtype=gpu.int16, device="cuda")
    st = gpu.cuda.current_stream()
    print(f"batch kernel={kernel} eager", flush=True)
    b.run(out.data_ptr(), st.cuda_stream, kernel, graph=False)
    gpu.cuda.synchronize()
    def squares(flat_all):
        return [cgh.upper_to_square(flat_all[b.offset(d): b.offset(d) + cgh.upper_elems(X.shape[0], S)],
                                    X.shape[0], S) for d, X in enumerate(Xs)]

    eager = squares(out.cpu().numpy().view(np.uint16))
    for d, X in enumerate(Xs):
        assert np.array_equal(eager[d], oracle.brute_full(X)), d
    side = gpu.cuda.Stream()  # graphs are captured on a non-default stream
    for rep in range(3):
        out.fill_(0x1234)
        gpu.cuda.synchronize()
        print(f"batch kernel={kernel} graph replay {rep}", flush=True)
        b.run(out.data_ptr(), side.cuda_stream, kernel, graph=True)
        gpu.cuda.synchronize()
        got = squares(out.cpu().numpy().view(np.uint16))
        assert all(np.array_equal(x, y) for x, y in zip(got, eager)), rep
    b.close()
    for s in sets:
        s.close()


def test_batch_autotune_then_exact(cgh, gpu, workloads):
    """The batch autotuner times every tile rule and product kernel over the
    whole batch (each candidate's output checked against the first's inside
    the call); the batch then runs the chosen rule under CGH_KERNEL_AUTO,
    eager and from a graph, every dataset equal to the oracle."""
    shapes = [(300, 97, 21), (2100, 300, 22), (4500, 200, 23), (9000, 120, 24)]
    S = 1024
    sets, Xs = [], []
    for n, L, seed in shapes:
        X = cgh.dense_ranks(workloads.zipf_iid(n, L, seed))
        Xs.append(X)
        sets.append(cgh.Dataset(cgh.Encoded(X, 16, 256)))
    b = cgh.Batch(sets, S)
    out = gpu.zeros((b.out_elems(),), dtype=gpu.int16, device="cuda")
    st = gpu.cuda.current_stream()
    print("batch autotune", flush=True)
    th, k, ms = b.autotune(out.data_ptr(), st.cuda_stream)
    assert th in (0, 2048, 4096, 8192, 16384, 2**31 - 1) and k in (1, 2) and ms > 0
    assert [b.tile(d) for d in range(len(shapes))] == [256 if n >= th else 128 for n, _, _ in shapes]
    want = [oracle.brute_full(X) for X in Xs]
    side = gpu.cuda.Stream()
    for graph, stream in ((False, st), (True, side), (True, side)):
        out.fill_(0x2b2b)
        gpu.cuda.synchronize()
        print(f"batch tuned graph={graph}", flush=True)
        b.run(out.data_ptr(), stream.cuda_stream, 0, graph=graph)
        gpu.cuda.synchronize()
        flat = out.cpu().numpy().view(np.uint16)
        for d, X in enumerate(Xs):
            n = X.shape[0]
            got = cgh.upper_to_square(flat[b.offset(d): b.offset(d) + cgh.upper_elems(n, S)], n, S)
            assert np.array_equal(got, want[d]), (graph, d)
    b.close()
    for s in sets:
        s.close()


def test_batch_with_caps(cgh, gpu, workloads):
    """A batch of datasets with different -X caps (none, 5, 40): every
    dataset's cells are min(H, cap), eager and replayed from the graph."""
    shapes = [(900, 300, 11, 0), (1300, 500, 12, 5), (600, 800, 13, 40)]
    S = 512
    sets, Xs, caps = [], [], []
    for n, L, seed, cap in shapes:
        X = cgh.dense_ranks(workloads.zipf_iid(n, L, seed))
        ds = cgh.Dataset(cgh.Encoded(X, 64, 256))
        if cap:
            ds.set_cap(cap)
        sets.append(ds)
        Xs.append(X)
        caps.append(cap)
    b = cgh.Batch(sets, S)
    out = gpu.zeros(b.out_elems(), dtype=gpu.int16, device="cuda")
    side = gpu.cuda.Stream()
    for graph in (False, True):
        out.fill_(0x2222)
        gpu.cuda.synchronize()
        print(f"batch with caps graph={graph}", flush=True)
        b.run(out.data_ptr(), side.cuda_stream, 0, graph=graph)
        gpu.cuda.synchronize()
        flat = out.cpu().numpy().view(np.uint16)
        for d, (X, cap) in enumerate(zip(Xs, caps)):
            n = X.shape[0]
            H = cgh.upper_to_square(flat[b.offset(d): b.offset(d) + cgh.upper_elems(n, S)], n, S)
            ref = oracle.brute_full(X)
            assert np.array_equal(H, np.minimum(ref, cap) if cap else ref), (d, graph)
    b.close()
    for ds in sets:
        ds.close()


def test_batch_with_a_dataset_beyond_u16_rows(cgh, gpu):
    """A batch holding a 70,000-row dataset (u32 member lists, wide rows in
    two LDS chunks) beside a small one: the batched launch equals the
    per-dataset launches, and sampled rows equal the oracle."""
    rng = np.random.default_rng(19)
    big = rng.integers(1, 300, size=(70000, 24)).astype(np.uint16)
    big[rng.random(big.shape) < 0.05] = 0
    small = cgh.dense_ranks(np.random.default_rng(3).integers(1, 40, size=(900, 60)).astype(np.uint64))
    S = 4096
    sets = [cgh.Dataset(cgh.Encoded(big, 4096, 256)), cgh.Dataset(cgh.Encoded(small, 64, 256))]
    b = cgh.Batch(sets, S)
    out = gpu.zeros(b.out_elems(), dtype=gpu.int16, device="cuda")
    print("batch with a 70000-row dataset", flush=True)
    b.run(out.data_ptr(), gpu.cuda.current_stream().cuda_stream, 0, graph=False)
    plain = gpu.zeros_like(out)
    for d, ds in enumerate(sets):
        n = (big if d == 0 else small).shape[0]
        view = plain[b.offset(d): b.offset(d) + cgh.upper_elems(n, S)]
        ds.upper(S, view.data_ptr(), gpu.cuda.current_stream().cuda_stream, 0)
    gpu.cuda.synchronize()
    for d, X in enumerate((big, small)):
        n = X.shape[0]
        seg = slice(b.offset(d), b.offset(d) + cgh.upper_elems(n, S))
        assert bool(gpu.equal(out[seg], plain[seg])), d
        for i in ([0, 4462, 4463, 35000, 69998] if d == 0 else [0, 450, 898]):
            s = i // S
            r0 = s * S
            off = b.offset(d) + cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
            got = out[off + (i + 1 - r0): off + (n - r0)].cpu().numpy().view(np.uint16)
            assert np.array_equal(got, oracle.brute_rows(X, i, i + 1)[0, i + 1:]), (d, i)
    b.close()
    for ds in sets:
        ds.close()


def test_rare_pass_many_chunks(cgh, gpu):
    """n = 70,000 > 65,536: member lists stored as u32, and the rows with more
    than 65,536 cells right of the diagonal (i < 4,463) split into two LDS
    chunks, each clipping every partner range by binary search; rows on both
    sides of that split and of the stripe edges checked against the oracle."""
    rng = np.random.default_rng(12)
    n, L = 70000, 24
    X = rng.integers(1, 300, size=(n, L)).astype(np.uint16)   # ~230 copies per allele: rare at tau=4096
    X[rng.random((n, L)) < 0.05] = 0
    S = 4096
    enc = cgh.Encoded(X, 4096, 256)
    assert enc.enc.n_rare > 10**8
    ds = cgh.Dataset(enc)
    out = gpu.zeros(cgh.upper_elems(n, S), dtype=gpu.int16, device="cuda")
    print("rare many chunks, u32 members", flush=True)
    ds.upper(S, out.data_ptr(), gpu.cuda.current_stream().cuda_stream)
    gpu.cuda.synchronize()
    ds.close()
    for i in [0, 1, 2, 4095, 4096, 4461, 4462, 4463, 4464, 8191, 40000, 69998]:
        s = i // S
        r0 = s * S
        w = n - r0
        off = cgh.stripe_offset(n, S, s) + (i - r0) * w
        got = out[off + (i + 1 - r0): off + w].cpu().numpy().view(np.uint16)
        ref = oracle.brute_rows(X, i, i + 1)[0, i + 1:]
        assert np.array_equal(got, ref), i


def upper_checksum(cgh, flat, n, S):
    """tensorcore/tc_hamming.cu:241's checksum over the strict upper triangle
    held in the stripe layout (row i's cells j > i are one contiguous run)."""
    h = 0
    for s, r0 in enumerate(range(0, n, S)):
        off = cgh.stripe_offset(n, S, s)
        w = n - r0
        for r in range(min(S, n - r0)):
            base = off + r * w
            h = oracle.checksum_cells(flat[base + r + 1: base + w], h)
    return h


def test_rare_flat_walk_one_row_items(cgh, gpu, workloads):
    """Rows longer than half an LDS chunk (n = 33,000 > 32,768): every rare
    work item is one row, which takes the flat walk (interior member words
    laid end to end over the batch, edge words per range), as on the
    63,005-row benchmark set.  Long-tail allele counts give ranges from one
    member to thousands.  The whole upper triangle, chained into
    tc_hamming.cu:241's checksum, equals the oracle's at every tau (from
    nearly all-frequent to all-rare) and both MFMA product forms; rows at the
    stripe edges equal the oracle's rows, also under a cap."""
    n, L, S = 33000, 40, 4096
    X = cgh.dense_ranks(workloads.long_tail(n, L, seed=31))
    print("flat walk: oracle checksum", flush=True)
    want = oracle.checksum_upper(X)
    rows = [0, 1, 4095, 4096, n // 2, n - 2]
    ref = {i: oracle.brute_rows(X, i, i + 1)[0, i + 1:] for i in rows}
    out = gpu.empty(cgh.upper_elems(n, S), dtype=gpu.int16, device="cuda")
    st = gpu.cuda.current_stream()
    for tau, kernel, cap in ((2, 2, 0), (64, 1, 0), (1024, 2, 0), (10**6, 2, 0), (64, 2, 17)):
        print(f"flat walk: tau={tau} kernel={kernel} cap={cap}", flush=True)
        enc = cgh.Encoded(X, tau, 256)
        assert tau < 10**6 or enc.enc.D == 0
        ds = cgh.Dataset(enc)
        if cap:
            ds.set_cap(cap)
        out.fill_(0x6b6b)
        ds.upper(S, out.data_ptr(), st.cuda_stream, kernel)
        gpu.cuda.synchronize()
        ds.close()
        flat = out.cpu().numpy().view(np.uint16)
        if not cap:
            assert upper_checksum(cgh, flat, n, S) == want, (tau, kernel)
        for i in rows:
            s = i // S
            r0 = s * S
            base = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
            got = flat[base + i + 1 - r0: base + n - r0]
            assert np.array_equal(got, np.minimum(ref[i], cap) if cap else ref[i]), (tau, kernel, cap, i)


@pytest.mark.parametrize("n", [20000, 5000])
def test_rare_flat_walk_multi_row_items(cgh, gpu, workloads, n):
    """Rows short enough that several share an LDS chunk (3 rows of 20,000,
    13 rows of 5,000 per work item).  With long member ranges (a mean of 160
    partners or more per range) such items take the flat walk over all their
    rows' entries at once: batches cross row boundaries, each range's row is
    searched from the rows' starts and carried with its member words.  With
    short ranges (tau 256, a mean near 40) they take the subgroup walk.  The
    whole upper triangle, chained into tc_hamming.cu:241's checksum, equals
    the oracle's either way; rows at the stripe and item edges equal the
    oracle's rows, also under a cap."""
    L, S = 1142, 4096
    X = cgh.dense_ranks(workloads.long_tail(n, L, seed=37))
    print(f"multi-row flat walk n={n}: oracle checksum", flush=True)
    want = oracle.checksum_upper(X)
    rows = [0, 1, 2, 3, 4095, 4096, n // 2, n - 2]
    ref = {i: oracle.brute_rows(X, i, i + 1)[0, i + 1:] for i in rows}
    out = gpu.empty(cgh.upper_elems(n, S), dtype=gpu.int16, device="cuda")
    st = gpu.cuda.current_stream()
    for tau, flat, cap in ((256, False, 0), (2048, True, 0), (10**6, True, 0), (2048, True, 250)):
        enc = cgh.Encoded(X, tau, 256)
        mean = int(enc.enc.n_rare) / max(1, int(enc.enc.n_rentries))
        assert (mean >= 160) == flat, (tau, mean)  # which walk the items take
        print(f"multi-row flat walk n={n}: tau={tau} mean={mean:.0f} cap={cap}", flush=True)
        ds = cgh.Dataset(enc)
        if cap:
            ds.set_cap(cap)
        out.fill_(0x6b6b)
        ds.upper(S, out.data_ptr(), st.cuda_stream, 2)
        gpu.cuda.synchronize()
        ds.close()
        flat_out = out.cpu().numpy().view(np.uint16)
        if not cap:
            assert upper_checksum(cgh, flat_out, n, S) == want, (n, tau)
        for i in rows:
            s = i // S
            r0 = s * S
            base = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
            got = flat_out[base + i + 1 - r0: base + n - r0]
            assert np.array_equal(got, np.minimum(ref[i], cap) if cap else ref[i]), (n, tau, cap, i)


@pytest.mark.parametrize("cap", [4464, 65535, 65536, 70000])
def test_cap_at_and_beyond_u16(cgh, gpu, cap):
    """-X at and past the u16 range on rows that the rare pass finishes (tau
    4096: every allele of these 300 rows is rare): a cap from 65,535 up
    changes nothing, since every distance is <= L, whatever the cap is modulo
    2^16 (70,000 = 65,536 + 4,464).  Dataset path, batch path and the CLI."""
    rng = np.random.default_rng(21)
    n, L, S = 300, 5000, 128
    X = rng.integers(1, 21, size=(n, L)).astype(np.uint16)   # 20 alleles: H ~ 0.95 L > 4,464
    X[rng.random((n, L)) < 0.02] = 0
    ref = oracle.brute_full(X)
    assert ref.max() > 4464
    want = np.minimum(ref, cap) if cap < 65535 else ref
    print(f"cap {cap}: dataset", flush=True)
    H, _, enc = gpu_upper(cgh, gpu, X, 4096, S, cap=cap)
    assert enc.enc.n_rare > 0
    assert np.array_equal(H, want)
    ds = cgh.Dataset(cgh.Encoded(X, 4096, 256))
    ds.set_cap(cap)
    b = cgh.Batch([ds], S)
    out = gpu.zeros(b.out_elems(), dtype=gpu.int16, device="cuda")
    print(f"cap {cap}: batch", flush=True)
    b.run(out.data_ptr(), gpu.cuda.current_stream().cuda_stream, 0, graph=False)
    gpu.cuda.synchronize()
    Hb = cgh.upper_to_square(out.cpu().numpy().view(np.uint16)[: cgh.upper_elems(n, S)], n, S)
    assert np.array_equal(Hb, want)
    b.close()
    ds.close()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "in.tsv")
        text = "FILE\t" + "\t".join(f"L{k}" for k in range(L)) + "\n"
        text += "".join(f"S{i}\t" + "\t".join(str(v) if v else "-" for v in r) + "\n" for i, r in enumerate(X))
        with open(path, "w") as f:
            f.write(text)
        print(f"cap {cap}: cli", flush=True)
        out = subprocess.run([cgh.CLI_PATH, "-s", "-i", path, "--tau", "4096", "-X", str(cap), "--stripe", "64"],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        assert out.stdout == oracle.render_matrix([f"S{i}" for i in range(n)], ref, cap=cap)
