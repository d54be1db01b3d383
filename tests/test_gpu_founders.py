"""GPU: operands built on the device (cgh_dataset_build) and the dense-frequent
founder shape at full size.

The device-built dataset must give exactly the upload path's and the
oracle's distances.  At 100,000 x 5,000 (50 founder alleles per locus, 1%
fresh mutations, 1% missing) tau = 1 puts every allele, the 5 million
singleton mutations included, into the one-hot (D ~ 5.25 million columns,
66 GB of O bits on the device); the result must be bit-identical to the
tau chosen by the cost model, and its first and last rows must equal the
oracle's brute force (integer results: tolerance 0).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def run_upper(cgh, torch, ds, n, S, kernel=0):
    out = torch.zeros(cgh.upper_elems(n, S), dtype=torch.int16, device="cuda")
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    ds.upper(S, out.data_ptr(), st.cuda_stream, kernel)
    e1.record(st)
    torch.cuda.synchronize()
    print(f"  upper triangle in {e0.elapsed_time(e1):.1f} ms", flush=True)
    return out


def rows_of(cgh, flat, n, S, i0, i1):
    """Strict upper part of rows i0..i1-1 from the stripe layout."""
    parts = []
    for i in range(i0, i1):
        s = i // S
        r0 = s * S
        off = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
        parts.append(flat[off + (i + 1 - r0): off + (n - r0)])
    return np.concatenate(parts)


@pytest.mark.parametrize("tau", [1, 8, 64, 4096])
@pytest.mark.parametrize("kernel", [2, 3])
def test_build_matches_upload_and_oracle(cgh, gpu, workloads, tau, kernel):
    X = cgh.dense_ranks(workloads.small_zipf(1500, 700, seed=13))
    n, S = X.shape[0], 512
    print(f"dataset build tau={tau} kernel={kernel}", flush=True)
    a = run_upper(cgh, gpu, cgh.Dataset.build(X, tau), n, S, kernel)
    b = run_upper(cgh, gpu, cgh.Dataset(cgh.Encoded(X, tau, 256)), n, S, kernel)
    assert bool(gpu.equal(a, b))
    flat = a.cpu().numpy().view(np.uint16)
    assert np.array_equal(rows_of(cgh, flat, n, S, 0, n), oracle.upper_block(X, 0, n))


def test_founders_full_size_tau1_equals_auto(cgh, gpu, workloads):
    n, L, S = 100000, 5000, 4096
    print("founders 100000 x 5000: generate + rank", flush=True)
    X = cgh.dense_ranks(workloads.founders(n, L))
    out_bytes = cgh.upper_elems(n, S) * 2
    tau = cgh.choose_tau(X, out_bytes, cgh.free_vram(0))
    assert tau >= 2
    print(f"founders tau=1 (build on device)", flush=True)
    ds1 = cgh.Dataset.build(X, 1)
    h1 = run_upper(cgh, gpu, ds1, n, S)
    ds1.close()
    print(f"founders tau={tau}", flush=True)
    dsa = cgh.Dataset.build(X, tau)
    ha = run_upper(cgh, gpu, dsa, n, S)
    dsa.close()
    assert bool(gpu.equal(h1, ha))
    del ha
    flat = h1.cpu().numpy().view(np.uint16)
    print("founders oracle rows", flush=True)
    for i0, i1 in ((0, 6), (n // 2, n // 2 + 2), (n - 6, n)):
        assert np.array_equal(rows_of(cgh, flat, n, S, i0, i1), oracle.upper_block(X, i0, i1)), (i0, i1)
