"""GPU: resident database + continuous-batching query server.

Every query x database distance is checked against the oracle's brute force
over the database and query profiles ranked together, the top-10 against a
full sort by (distance, row), and every request's result must be identical
when served alone and when mixed with concurrent requests (tolerance 0:
integer results).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SIZES = [1, 5, 17, 64, 3, 130, 40, 9]


def make_sets(workloads, n_db=3000, L=200, seed=21):
    db_calls, q_calls = workloads.serving_sets(n_db, L, sum(SIZES), seed)
    X_all = oracle.dense_ranks(np.vstack([db_calls, q_calls]))
    ref = oracle.brute_rows(X_all, n_db, X_all.shape[0])[:, :n_db]
    return db_calls, q_calls, ref


def expected_top(ref_rows):
    out_d, out_i = [], []
    for row in ref_rows:
        order = np.lexsort((np.arange(row.size), row))[:10]
        out_d.append(row[order])
        out_i.append(order)
    return np.array(out_d, dtype=np.uint16), np.array(out_i, dtype=np.uint32)


@pytest.mark.parametrize("tau,slots,kernel", [(32, 4, 0), (4, 2, 1), (10**6, 3, 2), (16, 2, 3)])
def test_server_alone_and_mixed(cgh, gpu, workloads, tau, slots, kernel):
    db_calls, q_calls, ref = make_sets(workloads)
    db = cgh.Database(db_calls, tau)
    sv = cgh.Server(db, slots=slots, max_batch=256, kernel=kernel)
    bounds = np.cumsum([0] + SIZES)
    # mixed: everything submitted at once, batches overlap on the device
    print(f"server tau={tau} slots={sv.slots} kernel={kernel} mixed", flush=True)
    tickets = [sv.submit(q_calls[bounds[k]:bounds[k + 1]], want_dist=k % 2 == 0) for k in range(len(SIZES))]
    mixed = [sv.wait(t) for t in tickets]
    # alone: one request in flight at a time
    print(f"server tau={tau} alone", flush=True)
    alone = [sv.wait(sv.submit(q_calls[bounds[k]:bounds[k + 1]])) for k in range(len(SIZES))]
    for k in range(len(SIZES)):
        want = ref[bounds[k]:bounds[k + 1]]
        td, ti = expected_top(want)
        for res in (mixed[k], alone[k]):
            dist, top_d, top_i, lat = res
            assert (dist is None) == (res is mixed[k] and k % 2 == 1), k
            assert dist is None or np.array_equal(dist, want), k
            assert np.array_equal(top_d, td) and np.array_equal(top_i, ti), k
            assert lat > 0
    sv.close()
    db.close()


def test_server_admission_by_memory(cgh, gpu, workloads):
    """A budget that holds about one 64-query batch: larger requests are
    refused up front, the rest are admitted one batch at a time and still
    come back exact."""
    db_calls, q_calls, ref = make_sets(workloads, n_db=2000, L=120, seed=5)
    db = cgh.Database(db_calls, 16)
    per64 = 64 * 2000 * 2 + 256 * 0 + 64 * 120 * 8 + 64 * 60 + 64 * 2000  # generous bound on one 64-query batch
    sv = cgh.Server(db, slots=4, max_batch=256, hbm_budget=per64)
    with pytest.raises(cgh.CghError, match="admission budget"):
        sv.submit(q_calls[:130])
    bounds = np.cumsum([0] + SIZES)
    small = [k for k in range(len(SIZES)) if SIZES[k] <= 40]
    print("server admission", flush=True)
    tickets = {k: sv.submit(q_calls[bounds[k]:bounds[k + 1]]) for k in small}
    for k, t in tickets.items():
        dist, top_d, top_i, _ = sv.wait(t)
        assert np.array_equal(dist, ref[bounds[k]:bounds[k + 1]]), k
    sv.close()
    db.close()


def test_server_database_smaller_than_top_k(cgh, gpu, workloads):
    """7 database rows: the 10-entry lists end with (0xffff, 0xffffffff)."""
    db_calls, q_calls = workloads.serving_sets(7, 60, 3, seed=9)
    X_all = oracle.dense_ranks(np.vstack([db_calls, q_calls]))
    ref = oracle.brute_rows(X_all, 7, 10)[:, :7]
    db = cgh.Database(db_calls, 2)
    sv = cgh.Server(db, slots=1, max_batch=4)
    dist, top_d, top_i, _ = sv.wait(sv.submit(q_calls))
    assert np.array_equal(dist, ref)
    td, ti = expected_top(ref)
    assert np.array_equal(top_d[:, :7], td[:, :7]) and np.array_equal(top_i[:, :7], ti[:, :7])
    assert np.all(top_d[:, 7:] == 0xFFFF) and np.all(top_i[:, 7:] == 0xFFFFFFFF)
    sv.close()
    db.close()


@pytest.mark.parametrize("tau", [64, 10**6])
def test_server_database_beyond_u16_rows(cgh, gpu, workloads, tau):
    """A 70,000-row database: member lists as u32 and every query row split
    into two LDS chunks of the rare pass (partner ranges clipped per chunk);
    distances and top-10 against the oracle."""
    n_db, L, sizes = 70000, 40, [3, 20]
    db_calls, q_calls = workloads.serving_sets(n_db, L, sum(sizes), seed=31)
    X_all = oracle.dense_ranks(np.vstack([db_calls, q_calls]))
    ref = oracle.brute_rows(X_all, n_db, X_all.shape[0])[:, :n_db]
    db = cgh.Database(db_calls, tau)
    sv = cgh.Server(db, slots=2, max_batch=64)
    print(f"server 70000-row database tau={tau}", flush=True)
    bounds = np.cumsum([0] + sizes)
    tickets = [sv.submit(q_calls[bounds[k]:bounds[k + 1]]) for k in range(len(sizes))]
    for k, t in enumerate(tickets):
        dist, top_d, top_i, _ = sv.wait(t)
        want = ref[bounds[k]:bounds[k + 1]]
        td, ti = expected_top(want)
        assert np.array_equal(dist, want), k
        assert np.array_equal(top_d, td) and np.array_equal(top_i, ti), k
    sv.close()
    db.close()
