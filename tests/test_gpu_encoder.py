"""GPU: the encoder on the device (cgh_dataset_encode).

Its dense ranks must equal the host's (cgh_dense_ranks, itself pinned to the
oracle), its tau choice the host rule's, and the distances of the dataset it
builds the oracle's brute force, cell for cell (integer results: tolerance
0).  At the 63,005 x 1,142 long-tail shape the device-encoded dataset must
give the same upper triangle as the host-encoded one.
"""
import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def upper(cgh, torch, ds, n, S):
    out = torch.zeros(cgh.upper_elems(n, S), dtype=torch.int16, device="cuda")
    ds.upper(S, out.data_ptr(), torch.cuda.current_stream().cuda_stream, 0)
    torch.cuda.synchronize()
    return out


def strict_upper(cgh, flat, n, S):
    parts = []
    for i in range(n):
        s = i // S
        r0 = s * S
        off = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
        parts.append(flat[off + (i + 1 - r0): off + (n - r0)])
    return np.concatenate(parts)


def datasets(workloads):
    a = workloads.small_zipf(1400, 600, seed=17)                        # 32-bit ids
    b = workloads.hashed_ids(workloads.zipf_iid(1100, 900, 23), bits=64, seed=5)
    c, _ = workloads.serving_sets(900, 400, 1, seed=7)                 # fresh + hashed 64-bit ids
    return {"small_zipf": a, "zipf64": b, "serving": c}


@pytest.mark.parametrize("tau", [1, 4, 64, 0])
@pytest.mark.parametrize("name", ["small_zipf", "zipf64", "serving"])
def test_device_encoder_matches_host_and_oracle(cgh, gpu, workloads, name, tau):
    calls = datasets(workloads)[name]
    n, L = calls.shape
    S = 512
    out_bytes = cgh.upper_elems(n, S) * 2
    print(f"device encoder {name} tau={tau}", flush=True)
    ds, tau_used, Xd = cgh.Dataset.encode(calls, tau, out_bytes, want_ranks=True)
    X = cgh.dense_ranks(calls)
    assert np.array_equal(Xd, X)
    want_tau = cgh.choose_tau(X, out_bytes, cgh.free_vram(0)) if tau == 0 else min(tau, n + 1)
    assert tau_used == want_tau
    flat = upper(cgh, gpu, ds, n, S).cpu().numpy().view(np.uint16)
    assert np.array_equal(strict_upper(cgh, flat, n, S), oracle.upper_block(X, 0, n))


def test_device_encoder_distinct_limit(cgh, gpu):
    calls = np.arange(1, 70001, dtype=np.uint64).reshape(70000, 1)
    with pytest.raises(cgh.CghError, match="locus 0 has 70000 distinct alleles"):
        cgh.Dataset.encode(calls, 8)


def test_device_encoder_long_tail_full_size(cgh, gpu, workloads):
    n, L, S = 63005, 1142, 4096
    calls = workloads.long_tail(n, L, seed=2)
    out_bytes = cgh.upper_elems(n, S) * 2
    t = time.perf_counter()
    X = cgh.dense_ranks(calls)
    tau = cgh.choose_tau(X, out_bytes, cgh.free_vram(0))
    host = cgh.Dataset(cgh.Encoded(X, tau, 256))
    t_host = time.perf_counter() - t
    gpu.cuda.synchronize()
    t = time.perf_counter()
    dev, tau_d, _ = cgh.Dataset.encode(calls, 0, out_bytes)
    t_dev = time.perf_counter() - t
    print(f"long tail: host encode + upload {t_host:.2f} s, device encode {t_dev:.2f} s, tau {tau} / {tau_d}",
          flush=True)
    assert tau_d == tau
    a = upper(cgh, gpu, host, n, S)
    b = upper(cgh, gpu, dev, n, S)
    assert bool(gpu.equal(a, b))


def test_device_encoder_beyond_u16_rows(cgh, gpu):
    """n = 70,000 > 65,536: the device encoder writes u32 member lists (and
    the rare pass splits the widest rows into two chunks); ranks equal the
    host's and sampled rows equal the oracle's."""
    rng = np.random.default_rng(41)
    n, L, S = 70000, 24, 4096
    calls = rng.integers(1, 300, size=(n, L)).astype(np.uint64) * 1000003
    calls[rng.random((n, L)) < 0.05] = 0
    out_bytes = cgh.upper_elems(n, S) * 2
    print("device encoder 70000 rows, tau 4096", flush=True)
    ds, tau_used, Xd = cgh.Dataset.encode(calls, 4096, out_bytes, want_ranks=True)
    X = cgh.dense_ranks(calls)
    assert np.array_equal(Xd, X) and tau_used == 4096
    assert ds.info()["n_rare"] > 10**8
    out = gpu.zeros(cgh.upper_elems(n, S), dtype=gpu.int16, device="cuda")
    ds.upper(S, out.data_ptr(), gpu.cuda.current_stream().cuda_stream, 0)
    gpu.cuda.synchronize()
    for i in [0, 4462, 4463, 35000, 69998]:
        s = i // S
        r0 = s * S
        off = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
        got = out[off + (i + 1 - r0): off + (n - r0)].cpu().numpy().view(np.uint16)
        assert np.array_equal(got, oracle.brute_rows(X, i, i + 1)[0, i + 1:]), i
    ds.close()
