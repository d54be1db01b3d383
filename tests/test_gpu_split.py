"""GPU parity of the one-hot split along K (cgh_dataset_upload_split): O
stays in host memory and the product runs one pass per chunk, the first
storing P P^T - O_0 O_0^T and each later one subtracting O_c O_c^T from the
stored cells.  The result must equal the in-core dataset's cell for cell and
the oracle's brute force, for both MFMA forms, with and without a cap, for
chunks of one slab up to the whole one-hot.  Integer results: tolerance 0."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def upper(cgh, gpu, ds, n, S, kernel):
    out = gpu.full((cgh.upper_elems(n, S),), 0x5a5a, dtype=gpu.int16, device="cuda")
    ds.upper(S, out.data_ptr(), gpu.cuda.current_stream().cuda_stream, kernel)
    gpu.cuda.synchronize()
    return out.cpu().numpy().view(np.uint16)


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("slabs", [1, 3, 1000])
@pytest.mark.parametrize("L", [600, 650])
def test_split_one_hot_equals_in_core(cgh, gpu, workloads, kernel, slabs, L):
    """L = 650 at 1-slab chunks ends with a pass holding one valid word, so
    the plane after its last O plane falls inside the slab (the case that
    caught an O-only pass negating its value twice)."""
    n, S, tau = 3000, 512, 4
    X = cgh.dense_ranks(workloads.small_zipf(n, L, seed=21))
    enc = cgh.Encoded(X, tau, 256)
    n_pad = int(enc.enc.n_pad)
    o_bytes = 2 * slabs * 4 * n_pad * 16  # two buffers of `slabs` 512-bit slabs of every row
    print(f"split one-hot L={L} kernel={kernel} slabs={slabs} D={int(enc.enc.D)}", flush=True)
    ref_ds = cgh.Dataset(enc)
    want = upper(cgh, gpu, ref_ds, n, S, kernel)
    ref_ds.close()
    ds = cgh.Dataset(enc, o_bytes=o_bytes)
    got = upper(cgh, gpu, ds, n, S, kernel)
    assert np.array_equal(got, want)
    H = cgh.upper_to_square(got, n, S)
    assert np.array_equal(H, oracle.brute_full(X))
    # the cap goes with the last pass (rows without rare partners) and with
    # the rare pass (the others)
    ds.set_cap(37)
    capped = cgh.upper_to_square(upper(cgh, gpu, ds, n, S, kernel), n, S)
    assert np.array_equal(capped, np.minimum(H, 37))
    ds.close()


def test_split_one_hot_refusals(cgh, gpu, workloads):
    """Entry points that need O on the device refuse a split dataset with an
    error (no launch)."""
    n, L = 700, 200
    X = cgh.dense_ranks(workloads.small_zipf(n, L, seed=22))
    enc = cgh.Encoded(X, 2, 256)
    ds = cgh.Dataset(enc, o_bytes=2 * 4 * int(enc.enc.n_pad) * 16)
    out = gpu.zeros((cgh.upper_elems(n, 256),), dtype=gpu.int16, device="cuda")
    with pytest.raises(cgh.CghError, match="MFMA"):
        ds.upper(256, out.data_ptr(), gpu.cuda.current_stream().cuda_stream, 3)
    with pytest.raises(cgh.CghError, match="host memory"):
        ds.stream(256)
    with pytest.raises(cgh.CghError, match="host memory"):
        ds.autotune(256)
    with pytest.raises(cgh.CghError, match="host memory"):
        cgh.Batch([ds], 256)
    with pytest.raises(cgh.CghError, match="less than"):
        cgh.Dataset(enc, o_bytes=1000)
    ds.close()
