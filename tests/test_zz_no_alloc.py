"""GPU: cgh_batch_run_noalloc, the batch entry point that allocates nothing,
against cgh_batch_run (eager) on the same datasets: every cell of every
dataset's layout identical (integer results, tolerance 0), on two calls in a
row (the second with fewer and smaller datasets, so stale table entries,
tile lists and counters of the first call would show), both with the
oracle's brute force on sampled rows; plus its size limits, rejected before
any device work.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def layout_cells(cgh, n, S):
    s = (n - 1) // S
    return cgh.stripe_offset(n, S, s) + (n - s * S) ** 2


def make_sets(cgh, workloads, shapes):
    Xs, sets = [], []
    for n, L, seed, tau in shapes:
        X = cgh.dense_ranks(workloads.batch_dataset(n, L, seed, seed))
        Xs.append(X)
        sets.append(cgh.Dataset(cgh.Encoded(X, tau, 256)))
    return Xs, sets


def check_call(cgh, torch, Xs, sets, S, kernel, caps=None):
    b = cgh.Batch(sets, S)
    st = torch.cuda.current_stream().cuda_stream
    ref = torch.full((b.out_elems(),), 0x1111, dtype=torch.int16, device="cuda")
    b.run(ref.data_ptr(), st, kernel, graph=False)
    out = torch.full((b.out_elems(),), 0x2222, dtype=torch.int16, device="cuda")
    print(f"noalloc: {len(sets)} datasets S={S} kernel={kernel}", flush=True)
    cgh.batch_run_noalloc(sets, S, out.data_ptr(), st, kernel)
    torch.cuda.synchronize()
    flat = out.cpu().numpy().view(np.uint16)
    for d, X in enumerate(Xs):
        n = X.shape[0]
        seg = slice(b.offset(d), b.offset(d) + layout_cells(cgh, n, S))
        assert bool(torch.equal(out[seg], ref[seg])), d
        for i in (0, n // 2, n - 2):
            s = i // S
            r0 = s * S
            base = b.offset(d) + cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
            want = oracle.brute_rows(X, i, i + 1)[0, i + 1:]
            if caps and caps[d]:
                want = np.minimum(want, caps[d])
            assert np.array_equal(flat[base + i + 1 - r0: base + n - r0], want), (d, i)
    b.close()


@pytest.mark.parametrize("kernel", [0, 1])
def test_noalloc_batch_equals_batch_run(cgh, gpu, workloads, kernel):
    S = 512
    first = [(9000, 700, 11, 256), (1300, 900, 12, 64), (5000, 300, 13, 4096), (700, 1200, 14, 8),
             (4100, 500, 15, 2), (2, 40, 16, 2)]
    second = [(1500, 400, 21, 32), (600, 800, 22, 1)]
    Xs, sets = make_sets(cgh, workloads, first)
    sets[1].set_cap(7)
    check_call(cgh, gpu, Xs, sets, S, kernel, caps=[0, 7, 0, 0, 0, 0])
    for ds in sets:
        ds.close()
    Xs, sets = make_sets(cgh, workloads, second)
    check_call(cgh, gpu, Xs, sets, S, kernel)
    for ds in sets:
        ds.close()


def test_noalloc_limits(cgh, gpu, workloads):
    Xs, sets = make_sets(cgh, workloads, [(600, 100, 31, 16)])
    out = gpu.zeros(cgh.upper_elems(600, 512) + 8, dtype=gpu.int16, device="cuda")
    with pytest.raises(cgh.CghError, match="multiple of 256"):
        cgh.batch_run_noalloc(sets, 100, out.data_ptr())
    with pytest.raises(cgh.CghError, match="datasets"):
        cgh.batch_run_noalloc(sets * 257, 512, out.data_ptr())
    with pytest.raises(cgh.CghError, match="16-byte aligned"):
        cgh.batch_run_noalloc(sets, 512, out.data_ptr() + 2)
    sets[0].close()
