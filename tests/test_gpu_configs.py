"""GPU parity at the benchmark configurations' own shapes.

* 63,005 x 1,142 long-tail profiles (the headline workload): the whole upper
  triangle computed on the device, chained into the checksum of
  tensorcore/tc_hamming.cu:241 (h = h * 1000003 + H over i ascending, j > i
  ascending, u64 wrap-around) and compared with the oracle's brute-force
  checksum of the same triangle; at the cost model's tau with the fp4
  product and at tau = 256 with the int8 product (the result may not depend
  on tau or on the product form).
* 64 independent datasets, n ~ U[500, 20,000], L ~ U[800, 3,000], ids
  relabelled to 32 bits in half of them and 64 bits in the rest: the batched
  persistent launch, eager and replayed from its HIP graph, equals the 64
  per-dataset launches cell for cell, and rows of every dataset equal the
  oracle.
* A resident 200,000 x 1,142 database with requests of 1-256 queries
  (log-uniform sizes): every request's distances and 10 nearest rows equal
  the oracle's brute force, served alone and mixed with the others.

Integer results: tolerance 0 everywhere.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def layout_cells(cgh, n, S):
    """Cells of the stripe layout of n rows (without the rounding to 16-byte
    words)."""
    s = (n - 1) // S
    return cgh.stripe_offset(n, S, s) + (n - s * S) ** 2


def triangle_checksum(cgh, flat, n, S):
    """tc_hamming.cu:241's checksum over the strict upper triangle held in
    the stripe layout: row i's cells j > i are one contiguous run."""
    h = 0
    for s, r0 in enumerate(range(0, n, S)):
        off = cgh.stripe_offset(n, S, s)
        w = n - r0
        for r in range(min(S, n - r0)):
            base = off + r * w
            h = oracle.checksum_cells(flat[base + r + 1: base + w], h)
    return h


@pytest.mark.timeout(900)
def test_long_tail_63005_full_triangle_checksum(cgh, gpu, workloads):
    n, L, S = 63005, 1142, 4096
    print("long tail 63005 x 1142: generate + rank", flush=True)
    X = cgh.dense_ranks(workloads.long_tail(n, L, seed=2))
    print("oracle checksum (brute force, every pair)", flush=True)
    want = oracle.checksum_upper(X)
    tau_auto = cgh.choose_tau(X, cgh.upper_elems(n, S) * 2, cgh.free_vram(0))
    assert tau_auto >= 2
    out = gpu.empty(cgh.upper_elems(n, S), dtype=gpu.int16, device="cuda")
    st = gpu.cuda.current_stream()
    for tau, kernel in ((tau_auto, 2), (256, 1)):
        print(f"long tail tau={tau} kernel={kernel}", flush=True)
        ds = cgh.Dataset(cgh.Encoded(X, tau, 256))
        out.fill_(0x7777)
        ds.upper(S, out.data_ptr(), st.cuda_stream, kernel)
        gpu.cuda.synchronize()
        ds.close()
        flat = out.cpu().numpy().view(np.uint16)
        assert triangle_checksum(cgh, flat, n, S) == want, (tau, kernel)
        # and one full row against the brute force (a checksum match is not
        # an accident, but a row pins the cells themselves)
        for i in (0, n // 2):
            s = i // S
            r0 = s * S
            base = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
            assert np.array_equal(flat[base + i + 1 - r0: base + n - r0], oracle.brute_rows(X, i, i + 1)[0, i + 1:])


@pytest.mark.timeout(900)
def test_batch_64_uneven_datasets(cgh, gpu, workloads):
    S = 4096
    shapes = workloads.batch_shapes(64, seed=4)
    print(f"batch64: generate + rank {len(shapes)} datasets", flush=True)

    def make(arg):
        d, (n, L, seed) = arg
        return cgh.dense_ranks(workloads.batch_dataset(n, L, seed, d))

    with ThreadPoolExecutor(8) as ex:
        Xs = list(ex.map(make, enumerate(shapes)))
    assert min(X.shape[0] for X in Xs) >= 500 and max(X.shape[0] for X in Xs) <= 20000
    sets = []
    for X in Xs:
        n = X.shape[0]
        tau = cgh.choose_tau(X, cgh.upper_elems(n, S) * 2, cgh.free_vram(0)) or 64
        sets.append(cgh.Dataset(cgh.Encoded(X, tau, 256)))
    b = cgh.Batch(sets, S)
    st = gpu.cuda.current_stream().cuda_stream
    plain = gpu.zeros(b.out_elems(), dtype=gpu.int16, device="cuda")
    print("batch64: per-dataset launches", flush=True)
    for d, ds in enumerate(sets):
        ds.upper(S, plain.data_ptr() + 2 * b.offset(d), st, 0)
    out = gpu.full((b.out_elems(),), 0x4141, dtype=gpu.int16, device="cuda")
    print("batch64: batched, eager", flush=True)
    b.run(out.data_ptr(), st, 0, graph=False)
    gpu.cuda.synchronize()

    def same_as_plain():
        # every cell of every dataset's layout (not the arena's alignment padding)
        for d, X in enumerate(Xs):
            seg = slice(b.offset(d), b.offset(d) + layout_cells(cgh, X.shape[0], S))
            if not bool(gpu.equal(out[seg], plain[seg])):
                return d
        return None

    assert same_as_plain() is None
    side = gpu.cuda.Stream()
    for rep in range(2):
        out.fill_(0x4242)
        gpu.cuda.synchronize()
        print(f"batch64: graph replay {rep}", flush=True)
        b.run(out.data_ptr(), side.cuda_stream, 0, graph=True)
        gpu.cuda.synchronize()
        assert same_as_plain() is None, rep
    del plain
    flat = out.cpu().numpy().view(np.uint16)
    print("batch64: oracle rows", flush=True)
    for d, X in enumerate(Xs):
        n = X.shape[0]
        for i in (0, n // 3, n - 2):
            s = i // S
            r0 = s * S
            base = b.offset(d) + cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
            got = flat[base + i + 1 - r0: base + n - r0]
            assert np.array_equal(got, oracle.brute_rows(X, i, i + 1)[0, i + 1:]), (d, i)
    b.close()
    for ds in sets:
        ds.close()


def expected_top(ref_rows):
    out_d, out_i = [], []
    for row in ref_rows:
        order = np.lexsort((np.arange(row.size), row))[:10]
        out_d.append(row[order])
        out_i.append(order)
    return np.array(out_d, dtype=np.uint16), np.array(out_i, dtype=np.uint32)


@pytest.mark.timeout(900)
def test_server_200k_database(cgh, gpu, workloads):
    n_db, L, n_req = 200000, 1142, 24
    rng = np.random.default_rng(8)
    sizes = np.minimum(256, np.floor(np.exp(rng.uniform(0, np.log(257), size=n_req)))).astype(int)
    sizes[:2] = (256, 1)  # both ends of the request range
    print(f"server 200k: generate {n_db} + {int(sizes.sum())} profiles", flush=True)
    calls = workloads.zipf_iid(n_db + int(sizes.sum()), L, 9, s=1.2, a_min=5, a_max=3000, miss=0.02)
    X_all = oracle.dense_ranks(calls)
    print("server 200k: oracle distances of every query", flush=True)
    ref = oracle.brute_rows(X_all, n_db, X_all.shape[0])[:, :n_db]
    del X_all
    tau = cgh.choose_tau(cgh.dense_ranks(calls[:n_db]), 0, cgh.free_vram(0)) or 1024
    print(f"server 200k: database at tau={tau}", flush=True)
    db = cgh.Database(calls[:n_db], tau)
    q_calls = calls[n_db:]
    del calls
    sv = cgh.Server(db, slots=4, max_batch=1024)
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    print("server 200k: mixed", flush=True)
    tickets = [sv.submit(q_calls[bounds[k]:bounds[k + 1]], want_dist=True) for k in range(n_req)]
    mixed = [sv.wait(t) for t in tickets]
    print("server 200k: alone", flush=True)
    alone = [sv.wait(sv.submit(q_calls[bounds[k]:bounds[k + 1]], want_dist=True)) for k in range(n_req)]
    for k in range(n_req):
        want = ref[bounds[k]:bounds[k + 1]]
        td, ti = expected_top(want)
        for res in (mixed[k], alone[k]):
            dist, top_d, top_i, _ = res
            assert np.array_equal(dist, want), k
            assert np.array_equal(top_d, td) and np.array_equal(top_i, ti), k
    sv.close()
    db.close()
