#!/usr/bin/env python3
"""Benchmark: exact u16 allelic Hamming distances on one MI355X.

Headline: the whole upper triangle of the long-tail 63,005 x 1,142 profile
set (the metric's configuration), operands resident in HBM.  A step is one
pass of the hot path: the MFMA binary product with its fused u16 epilogue
(every stripe in one launch), then the row-local rare-pair pass, output
resident in HBM.  The same JSON line carries
  roofline / rooflines   the product (MFMA-bound) and the rare pass (bytes
                         that must reach HBM; its measured limiter named)
  batch                  64 uneven datasets: one batched persistent launch
                         (HIP graph replay) against one launch pair per
                         dataset, identical output checked
  serve                  a resident 200,000 x 1,142 database answering a
                         Poisson stream of 1-256-query requests: p50 / p99
                         latency and saturated Gpairs/s of the continuous-
                         batching server against plain per-request service,
                         identical results checked
  cpu_baseline           the oracle's direct compare built for this host's
                         CPU, timed on the 2,000 x 1,142 set in full and on
                         a sample of the headline set
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cgmlst-hamming-gpu_amd")

N, L, S = 63005, 1142, 4096
# MI355X published peaks per product kernel (CGH_KERNEL_*), in the units of
# the algorithmic count 2*(D+L) ops per pair.  The VALU form has no published
# popcount figure: its peak is derived from the published 157.3 TFLOP/s vector
# fp32 rate (78.6 T lane-instructions/s), one AND + one v_bcnt per 32 bit
# pairs = 32 ops per lane-instruction.
PEAKS = {1: (5000.0, "TOP/s", "int8 (0/1 operands, int32 accumulate)"),
         2: (10100.0, "TFLOP/s", "fp4 e2m1 (bits as 0 or 0.5/1/2 against reciprocal partners: every product 0 or 1; "
                                 "fp32 accumulate of integers <= L: exact)"),
         3: (2517.0, "TOP/s", "u32 (VALU AND + popcount, u32 accumulate)")}
MEASURED_PEAK_KEY = {1: "int8_mfma_tops", 2: "fp4_mfma_tflops", 3: "valu_popcount_tops"}
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_package():
    name = "cgmlst_hamming_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def kernel_source_sha():
    """sha256 of the kernel source the library was built from."""
    import hashlib
    with open(os.path.join(PKG_DIR, "csrc", "hamming.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def summarize_counters(dirs, tau, kern, out_path):
    """FETCH_SIZE / WRITE_SIZE passes of the headline and of --copy-calibration
    (rocprofv3 --pmc ... --output-format csv, one directory each) -> HBM bytes
    per launch of the product and rare kernels, scaled by what the same
    counter reads for the 1 GiB copy, written as JSON for the bench line."""
    import csv
    from collections import defaultdict

    def per_dispatch(d, counter):
        vals = defaultdict(float)
        names = {}
        with open(os.path.join(d, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    vals[r["Dispatch_Id"]] += float(r["Counter_Value"])
                    names[r["Dispatch_Id"]] = r["Kernel_Name"]
        return [(names[k], v) for k, v in vals.items()]

    def kernel_avg(rows, key):
        v = [x for name, x in rows if key in name]
        return sum(v) / len(v) if v else None

    fetch, write, cfetch, cwrite = dirs
    copy_bytes = 1 << 30
    cf = max(x for _, x in per_dispatch(cfetch, "FETCH_SIZE"))
    cw = max(x for _, x in per_dispatch(cwrite, "WRITE_SIZE"))
    ff, wf = copy_bytes / 1024 / cf, copy_bytes / 1024 / cw
    rec = {"workload": "long_tail", "n": N, "L": L, "tau": tau, "stripe": S, "product_kernel": kern,
           "kernel_source_sha256": kernel_source_sha(),
           "calibration": {"copy_bytes": copy_bytes, "FETCH_SIZE_KiB": cf, "WRITE_SIZE_KiB": cw,
                           "fetch_factor": ff, "write_factor": wf},
           "kernels": {}}
    for label, key in (("product", "valu_tile" if kern == 3 else "mfma_tile"), ("rare", "sub_rare_rows")):
        fk = kernel_avg(per_dispatch(fetch, "FETCH_SIZE"), key)
        wk = kernel_avg(per_dispatch(write, "WRITE_SIZE"), key)
        rec["kernels"][label] = {"name_match": key, "FETCH_SIZE_KiB": fk, "WRITE_SIZE_KiB": wk,
                                 "fetch_bytes": fk * 1024 * ff, "write_bytes": wk * 1024 * wf,
                                 "traffic_bytes": fk * 1024 * ff + wk * 1024 * wf}
    with open(out_path, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec["kernels"]), flush=True)


def counter_traffic(tau, kern):
    """HBM-side bytes per launch of the product and rare kernels from the
    committed counter passes, when they cover this configuration and this
    kernel source; else None for both."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_long_tail_63005x1142.json")
    if not os.path.exists(tpath):
        return {"product": None, "rare": None}
    t_rec = json.load(open(tpath))
    if ((t_rec["n"], t_rec["L"], t_rec["tau"], t_rec["stripe"], t_rec.get("product_kernel")) != (N, L, tau, S, kern)
            or t_rec.get("kernel_source_sha256") != kernel_source_sha()):
        return {"product": None, "rare": None}
    return {k: t_rec["kernels"][k]["traffic_bytes"] for k in ("product", "rare")}


def headline(args, cgh, wl, torch):
    """The 63,005 x 1,142 long-tail set: product + rare pass per step."""
    t = time.perf_counter()
    calls = wl.long_tail(N, L, seed=2)
    X = cgh.dense_ranks(calls)
    del calls
    log(f"workload long_tail {N}x{L} generated + ranked in {time.perf_counter() - t:.1f}s")
    tau = args.tau or cgh.choose_tau(X, cgh.upper_elems(N, S) * 2, cgh.free_vram(0), log=log)
    t = time.perf_counter()
    enc = cgh.Encoded(X, tau, 256)
    e = enc.enc
    D, R = int(e.D), int(e.n_rare)
    members, entries = int(e.n_members), int(e.n_rentries)
    log(f"encoded tau={tau} D={D} rare_pairs={R} kw1={e.kw1} kw2={e.kw2} in {time.perf_counter() - t:.1f}s")
    ds = cgh.Dataset(enc, 0)
    del enc
    out = torch.empty(cgh.upper_elems(N, S), dtype=torch.int16, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step(ev=None):
        if ev:
            ev[0].record(stream)
        ds.phase(S, out.data_ptr(), cgh.PHASE_PRODUCT, sp, args.kernel)
        if ev:
            ev[1].record(stream)
        ds.phase(S, out.data_ptr(), cgh.PHASE_RARE, sp)
        if ev:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prod_ms = sum(evs[k][0].elapsed_time(evs[k][1]) for k in range(args.steps)) / args.steps
    rare_ms = sum(evs[k][1].elapsed_time(evs[k][2]) for k in range(args.steps)) / args.steps
    ms_step = wall * 1e3 / args.steps
    pairs = N * (N - 1) // 2
    kern = 2 if args.kernel == 0 else args.kernel  # the automatic choice is fp4
    peak, punit, dtype = PEAKS[kern]
    log(f"step {ms_step:.2f} ms (product {prod_ms:.2f} ms, rare {rare_ms:.2f} ms)")

    measured = cgh.measure_peaks(0)
    memory = cgh.measure_memory(0)
    log("measured peaks: " + ", ".join(f"{k} {v:.0f}" for k, v in measured.items()) + "; memory: " +
        ", ".join(f"{k} {v:.0f}" for k, v in memory.items()))
    traffic = counter_traffic(tau, kern)

    # product: 2 (D + L) ops per pair (a multiply-add per bit pair)
    achieved = 2.0 * pairs * (D + L) / (prod_ms * 1e-3) / 1e12
    # rare pass: the bytes that must reach HBM are the output cells read and
    # written once (4 B per pair) and the member lists and range index read
    # once; the 2 B member index read per rare pair comes from the caches
    m_bytes = 2 if N <= 65535 else 4
    rare_hbm = 4.0 * pairs + m_bytes * members + 8.0 * entries + 4.0 * (N + 1)
    rare_gbs = rare_hbm / (rare_ms * 1e-3) / 1e9
    rooflines = {
        "product": {"bound": "valu" if kern == 3 else "mfma", "achieved": achieved, "peak": peak, "unit": punit,
                    "frac": achieved / peak, "traffic": traffic["product"],
                    "kernel": ("valu_tile" if kern == 3 else "mfma_tile") +
                              " (binary product + u16 epilogue, one launch per step)",
                    "work": "2*(D+L) ops per pair (multiply-add per bit pair), pairs = n(n-1)/2, D unpadded",
                    "ms": prod_ms, "frac_of_measured": achieved / measured[MEASURED_PEAK_KEY[kern]]},
        "rare": {"bound": "hbm", "achieved": rare_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": rare_gbs / HBM_PEAK_GBS, "traffic": traffic["rare"],
                 "kernel": "sub_rare_rows (rare equal pairs subtracted in LDS, one launch per step)",
                 "work": "4 B per pair (u16 cell read + write) + member lists and range index once "
                         f"({members} members, {entries} ranges); the {R} member reads (2 B each) hit the caches",
                 "ms": rare_ms, "frac_of_measured_copy": rare_gbs / memory["copy_gbs"],
                 "rare_pairs_per_s": R / (rare_ms * 1e-3),
                 "limiter": "the walk (about 3.4 ms with staging and write-back switched off; ~11 vector "
                            "instructions per rare pair) plus the staging and write-back of items too wide "
                            "to be copied into LDS during the walk before them; see "
                            "profiles/sub_rare_rows_counters_*.md"},
    }
    if traffic["rare"]:  # L2 <-> fabric bytes the counters saw, against the cache-resident read rate
        rooflines["rare"]["l2_fabric"] = {"achieved": traffic["rare"] / (rare_ms * 1e-3) / 1e9,
                                          "ceiling": memory["cache_read_gbs"], "unit": "GB/s"}
    dominant = "product" if prod_ms >= rare_ms else "rare"
    rec = {
        "metric": "Gpairs/s exact H at 63k x 1142 (u16 upper triangle, device-resident)",
        "value": pairs / (ms_step * 1e-3) / 1e9,
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded long-tail Zipf(1.1) allele calls, Beta(0.5,15) per-sample missing)",
        "config": {"workload": "long_tail", "n": N, "L": L, "tau": tau, "stripe": S, "layout": "upper-tri u16",
                   "one_hot_cols": D, "rare_pairs": R, "tile": ds.tile(), "kernel": kern},
        "roofline": rooflines[dominant],
        "rooflines": rooflines,
        "phases_ms": {"product": prod_ms, "rare": rare_ms},
        "measured_peaks": dict(measured, **{"read_gbs": memory["read_gbs"], "cache_read_gbs": memory["cache_read_gbs"],
                                            "copy_gbs": max(measured["copy_gbs"], memory["copy_gbs"])}),
        "frac_of_measured_peak": achieved / measured[MEASURED_PEAK_KEY[kern]],
    }
    flat = out.cpu().numpy().view("uint16")
    del out
    ds.close()
    return rec, X, flat


def batch_record(args, cgh, wl, torch):
    """64 independent datasets (n ~ U[500, 20000], L ~ U[800, 3000], Zipf(1.2)
    long tail, 2% missing, ids relabelled to 32 / 64 bits), each with its own
    cost-model tau: the batched persistent launch replayed from its HIP graph
    against one launch pair per dataset (each with its autotuned kernel and
    tile), the same output arena, outputs compared cell for cell."""
    from collections import Counter
    from concurrent.futures import ThreadPoolExecutor
    Sb = 4096
    shapes = wl.batch_shapes(64, seed=4)
    t = time.perf_counter()

    def make(arg):
        d, (n, Ln, seed) = arg
        return cgh.dense_ranks(wl.batch_dataset(n, Ln, seed, d))

    with ThreadPoolExecutor(8) as ex:
        Xs = list(ex.map(make, enumerate(shapes)))
    sets, taus, tuned, pairs = [], [], [], 0
    for X in Xs:
        n = X.shape[0]
        tau = cgh.choose_tau(X, cgh.upper_elems(n, Sb) * 2, cgh.free_vram(0)) or 64
        ds = cgh.Dataset(cgh.Encoded(X, tau, 256), 0)
        tuned.append(ds.autotune(Sb)[:2])
        sets.append(ds)
        taus.append(tau)
        pairs += n * (n - 1) // 2
    del Xs
    log(f"batch: {len(shapes)} datasets, {pairs} pairs, generated + encoded + tuned in "
        f"{time.perf_counter() - t:.1f}s; taus {sorted(set(taus))}")
    b = cgh.Batch(sets, Sb)
    out = torch.zeros(b.out_elems(), dtype=torch.int16, device="cuda")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    base = out.data_ptr()

    def plain():
        for d, ds in enumerate(sets):
            ds.upper(Sb, base + 2 * b.offset(d), sp, 0)

    def batched():
        b.run(base, sp, 0, graph=True)

    def timed(fn, reps):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    with torch.cuda.stream(stream):
        plain()
        torch.cuda.synchronize()
        ref = out.clone()
        out.zero_()
        torch.cuda.synchronize()
        batched()
        torch.cuda.synchronize()
    same = True
    for d, ds in enumerate(sets):
        n = ds.n
        cells = cgh.stripe_offset(n, Sb, (n - 1) // Sb) + (n - ((n - 1) // Sb) * Sb) ** 2
        seg = slice(b.offset(d), b.offset(d) + cells)
        same = same and bool(torch.equal(out[seg], ref[seg]))
    del ref
    reps = max(3, args.steps // 4)
    ms_plain = timed(plain, reps)
    ms_batch = timed(batched, reps)
    log(f"batch: plain {ms_plain:.2f} ms, batched {ms_batch:.2f} ms, identical output: {same}")
    b.close()
    for ds in sets:
        ds.close()
    return {
        "workload": "64 datasets, n ~ U[500, 20000], L ~ U[800, 3000], Zipf(1.2), 2% missing, 32/64-bit ids",
        "pairs": pairs, "stripe": Sb, "steps": reps,
        "batched": {"ms": ms_batch, "gpairs_s": pairs / (ms_batch * 1e-3) / 1e9,
                    "form": "mfma_tile_queue per tile size + one rare launch per dataset on 4 streams, HIP graph"},
        "plain": {"ms": ms_plain, "gpairs_s": pairs / (ms_plain * 1e-3) / 1e9,
                  "form": "cgh_dataset_upper per dataset (its autotuned kernel and tile), one stream"},
        "speedup": ms_plain / ms_batch, "identical_output": same,
        "tuned_kernel_tile": {f"{k}/{t}": c for (k, t), c in Counter(tuned).items()},
    }


def serve_record(args, cgh, wl, torch):
    """A resident 200,000 x 1,142 database (Zipf(1.2) long tail, 2% missing)
    and requests of 1-256 query profiles (log-uniform sizes).  The
    continuous-batching server (4 slots, batches of up to 1,024 queries)
    against plain per-request service (one request at a time: encode, upload,
    product, rare pass, top-10, download, then the next), each on the same
    Poisson stream at 50 req/s (latency from the scheduled arrival to the
    result) and saturated (every request at once).  The 10 nearest rows of
    every request must be identical in all runs."""
    import threading
    import numpy as np
    n_db, Ls, n_req, rate = 200000, 1142, args.requests, 50.0
    rng = np.random.default_rng(8)
    sizes = np.minimum(256, np.floor(np.exp(rng.uniform(0, np.log(257), size=n_req)))).astype(int)
    t = time.perf_counter()
    calls = wl.zipf_iid(n_db + int(sizes.sum()), Ls, 9, s=1.2, a_min=5, a_max=3000, miss=0.02)
    db_calls, q_calls = calls[:n_db], calls[n_db:]
    del calls
    tau = args.tau_serve or cgh.choose_tau(cgh.dense_ranks(db_calls), 0, cgh.free_vram(0)) or 1024
    db = cgh.Database(db_calls, tau)
    del db_calls
    log(f"serve: database {n_db}x{Ls} tau={db.tau} one-hot={db.D} resident in {time.perf_counter() - t:.1f}s; "
        f"{n_req} requests, {int(sizes.sum())} queries")
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    gaps = rng.exponential(1.0 / rate, size=n_req)
    arrivals = np.cumsum(gaps)

    def served_run(sv, at):
        """Requests submitted at their arrival times by a submitter thread,
        waited for in order; latency from arrival to result."""
        tickets = [None] * n_req
        t0 = time.perf_counter()

        def submitter():
            for k in range(n_req):
                d = t0 + at[k] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[k] = sv.submit(q_calls[bounds[k]:bounds[k + 1]], want_dist=False)

        th = threading.Thread(target=submitter)
        th.start()
        lats, tops = [], []
        for k in range(n_req):
            while tickets[k] is None:
                time.sleep(0.0002)
            _, td, ti, _ = sv.wait(tickets[k])
            lats.append(time.perf_counter() - (t0 + at[k]))
            tops.append((td, ti))
        th.join()
        return time.perf_counter() - t0, np.array(lats), tops

    def plain_run(sv, at):
        """One request at a time, each started at its arrival time or when
        the previous one is done, whichever is later."""
        t0 = time.perf_counter()
        lats, tops = [], []
        for k in range(len(at)):
            d = t0 + at[k] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            _, td, ti, _ = sv.wait(sv.submit(q_calls[bounds[k]:bounds[k + 1]], want_dist=False))
            lats.append(time.perf_counter() - (t0 + at[k]))
            tops.append((td, ti))
        return time.perf_counter() - t0, np.array(lats), tops

    zeros = np.zeros(n_req)
    pairs = int(sizes.sum()) * n_db
    # Python's cyclic garbage collector stays off during the timed runs: a
    # collection pause in the submitting or waiting thread would be counted
    # as a request's latency.
    import gc
    gc.collect()
    gc.disable()
    try:
        # each saturated run right after a warm-up burst (every slot and
        # kernel warm, clocks up), then the Poisson stream
        sv = cgh.Server(db, slots=4, max_batch=1024)
        served_run(sv, zeros)
        sat_wall, sat_lat, sat_tops = served_run(sv, zeros)
        s_wall, s_lat, s_tops = served_run(sv, arrivals)
        slots = sv.slots
        sv.close()
        pv = cgh.Server(db, slots=1, max_batch=256)
        plain_run(pv, zeros[:8])
        psat_wall, psat_lat, psat_tops = plain_run(pv, zeros)
        p_wall, p_lat, p_tops = plain_run(pv, arrivals)
        pv.close()
    finally:
        gc.enable()
    db.close()
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
               for run in (s_tops, p_tops, psat_tops) for a, b in zip(sat_tops, run))
    pct = lambda v, q: float(np.percentile(v, q) * 1e3)
    log(f"serve: served p50 {pct(s_lat, 50):.2f} ms p99 {pct(s_lat, 99):.2f} ms, saturated "
        f"{pairs / sat_wall / 1e9:.1f} Gpairs/s; plain p50 {pct(p_lat, 50):.2f} ms p99 {pct(p_lat, 99):.2f} ms, "
        f"saturated {pairs / psat_wall / 1e9:.1f} Gpairs/s; identical top-10: {same}")
    return {
        "workload": f"{n_db} x {Ls} resident database, {n_req} requests of 1-256 queries (log-uniform), "
                    f"{int(sizes.sum())} queries",
        "tau": int(db.tau), "rate_req_s": rate, "pairs": pairs,
        "p50_ms": pct(s_lat, 50), "p99_ms": pct(s_lat, 99),
        "served": {"form": f"continuous batching, {slots} slots, batches of up to 1024 queries",
                   "poisson": {"p50_ms": pct(s_lat, 50), "p99_ms": pct(s_lat, 99),
                               "gpairs_s": pairs / s_wall / 1e9},
                   "saturated": {"gpairs_s": pairs / sat_wall / 1e9, "p50_ms": pct(sat_lat, 50),
                                 "p99_ms": pct(sat_lat, 99)}},
        "plain": {"form": "one request at a time (encode, upload, product, rare pass, top-10, download)",
                  "poisson": {"p50_ms": pct(p_lat, 50), "p99_ms": pct(p_lat, 99), "gpairs_s": pairs / p_wall / 1e9},
                  "saturated": {"gpairs_s": pairs / psat_wall / 1e9, "p50_ms": pct(psat_lat, 50),
                                "p99_ms": pct(psat_lat, 99)}},
        "saturated_speedup": psat_wall / sat_wall,
        "identical_top10": same,
    }


def split_record(args, cgh, wl, torch):
    """The one-hot split along K (cgh_dataset_upload_split) on the 63k
    long-tail set at tau 64 (a 0.68 GB one-hot of 86k columns): O stays in
    pinned host memory and streams through device chunks of an eighth of it,
    one product pass per chunk; timed against the in-core product over the
    same operands, outputs compared cell for cell."""
    import numpy as np

    tau = 64
    t = time.perf_counter()
    X = cgh.dense_ranks(wl.long_tail(N, L, seed=2))
    enc = cgh.Encoded(X, tau, 256)
    del X
    e = enc.enc
    o_bytes_full = int(e.n_pad) * ((int(e.kw1) + 15) // 16 * 16) * 4
    o_bytes = max(o_bytes_full // 8, 2 * 4 * int(e.n_pad) * 16)
    log(f"split: tau={tau} D={int(e.D)} one-hot {o_bytes_full / 1e9:.2f} GB, device chunks {o_bytes / 1e9:.3f} GB "
        f"(encoded in {time.perf_counter() - t:.1f}s)")
    stream = torch.cuda.current_stream()
    out = torch.empty(cgh.upper_elems(N, S), dtype=torch.int16, device="cuda")
    res = {}
    flats = []
    for label, kw in (("in_core", {}), ("split", {"o_bytes": o_bytes})):
        ds = cgh.Dataset(enc, 0, **kw)
        ds.phase(S, out.data_ptr(), cgh.PHASE_PRODUCT, stream.cuda_stream)
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record(stream)
            ds.phase(S, out.data_ptr(), cgh.PHASE_PRODUCT, stream.cuda_stream)
            ev[1].record(stream)
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        flats.append(out.cpu())
        ds.close()
        res[label] = {"product_ms": float(np.median(ts))}
    chunk_planes = o_bytes // 2 // (int(e.n_pad) * 16) // 4 * 4
    planes = -(-int(e.D) // 128)
    rec = {"workload": f"long_tail {N} x {L}, tau {tau}, D {int(e.D)}", "one_hot_gb": o_bytes_full / 1e9,
           "device_chunk_gb": o_bytes / 1e9, "passes": -(-planes // chunk_planes),
           "in_core": res["in_core"], "split": res["split"],
           "slowdown": res["split"]["product_ms"] / res["in_core"]["product_ms"],
           "identical_output": bool(torch.equal(flats[0], flats[1]))}
    log(f"split: in-core product {res['in_core']['product_ms']:.1f} ms, split ({rec['passes']} passes) "
        f"{res['split']['product_ms']:.1f} ms, identical output: {rec['identical_output']}")
    return rec


def native_oracle():
    """oracle/oracle.c compiled for this host's CPU (-march=native, OpenMP)
    into a temporary directory, loaded as the oracle's library."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    d = tempfile.mkdtemp(prefix="cgh_oracle_native_")
    so = os.path.join(d, "liboracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-o", so,
                    os.path.join(ROOT, "oracle", "oracle.c")], check=True)
    oracle.use_library(so)
    return oracle


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(X, flat, cgh, wl, target_s=12.0):
    """The reference's only CPU distance code is the direct compare
    sum_k [a!=0][b!=0][a!=b] (check.py:39-44, tc_hamming.cu:244-248); the
    oracle's OpenMP C restatement, built for this CPU, timed on the
    2,000 x 1,142 set in full and on the first rows of the headline set
    against all later rows (those rows of the GPU output are checked)."""
    import numpy as np
    oracle = native_oracle()
    n = X.shape[0]
    # the 2,000 x 1,142 set, whole upper triangle, repeated to at least 1 s
    X1 = cgh.dense_ranks(wl.small_zipf(2000, 1142, seed=1))
    reps, t = 0, time.perf_counter()
    while True:
        oracle.upper_block(X1, 0, X1.shape[0])
        reps += 1
        if time.perf_counter() - t >= 1.0:
            break
    c1 = reps * (2000 * 1999 // 2) / (time.perf_counter() - t) / 1e9
    probe = 256
    t = time.perf_counter()
    oracle.upper_block(X, 0, probe)
    per_row = (time.perf_counter() - t) / probe
    rows = int(min(n // 2, max(probe, target_s / max(per_row, 1e-9))))
    t = time.perf_counter()
    ref = oracle.upper_block(X, 0, rows)
    dt = time.perf_counter() - t
    pairs = int(ref.size)
    parts = []
    for i in range(rows):
        s = i // S
        r0 = s * S
        off = cgh.stripe_offset(n, S, s) + (i - r0) * (n - r0)
        parts.append(flat[off + (i + 1 - r0): off + (n - r0)])
    exact = bool(np.array_equal(np.concatenate(parts), ref))
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {
        "value": pairs / dt / 1e9, "unit": "Gpairs/s", "cores": oracle.num_threads(), "kind": "port",
        "sample": f"rows 0..{rows - 1} of the 63,005 x 1,142 set vs all later rows ({pairs} pairs, {dt:.1f} s)",
        "config1_full": {"gpairs_s": c1, "workload": "2,000 x 1,142 (small_zipf), whole upper triangle",
                         "repetitions": reps},
        "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity": affinity,
        "build": "gcc -O3 -march=native -fopenmp (oracle/oracle.c, built on this host at run time)",
        "op": "OpenMP C direct compare sum_k [a!=0][b!=0][a!=b] over u16 ranks, 16 x 16 pair tiles",
        "sample_matches_gpu": exact,
    }


def run_cli(args, cgh, wl):
    """End to end through the drop-in command line: the 63,005 x 1,142
    long-tail calls as a chewBBACA-style TSV -> the full u16 matrix as TSV
    (to /dev/null: formatting and writes without a disk) and gzip-compressed
    (to a file).  Reports the wall time and the CLI's own phase lines."""
    import re
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    calls = wl.long_tail(N, L, seed=2)
    tmp = tempfile.mkdtemp(prefix="cgh_cli_")
    try:
        path = os.path.join(tmp, "calls.tsv")
        t = time.perf_counter()
        tbl = np.array([str(v) for v in range(int(calls.max()) + 1)], dtype=object)
        with open(path, "w") as f:
            f.write("FILE\t" + "\t".join(f"L{k}" for k in range(L)) + "\n")
            for i in range(N):
                f.write(f"S{i}\t" + "\t".join(tbl[calls[i]]) + "\n")
        log(f"cli: wrote {os.path.getsize(path) / 1e6:.0f} MB of TSV in {time.perf_counter() - t:.1f}s")
        runs = {}
        for name, extra in (("full_tsv", ["-o", "/dev/null"]), ("full_gzip", ["-z", "-o", os.path.join(tmp, "o.gz")])):
            t = time.perf_counter()
            out = subprocess.run([cgh.CLI_PATH, "-i", path, *extra], capture_output=True, text=True, timeout=600)
            wall = time.perf_counter() - t
            if out.returncode:
                raise RuntimeError(out.stderr)
            phases = {}
            for key, pat in (("load", r"Loaded .* in ([0-9.]+)s"), ("encode", r"Encoded in ([0-9.]+)s"),
                             ("compute_and_stream", r"Computed .* in ([0-9.]+)s"),
                             ("write_tail", r"Wrote output in ([0-9.]+)s")):
                m = re.search(pat, out.stderr)
                if m:
                    phases[key] = float(m.group(1))
            runs[name] = {"wall_s": wall, "phases_s": phases}
            if name == "full_gzip":
                runs[name]["gzip_bytes"] = os.path.getsize(os.path.join(tmp, "o.gz"))
            log(f"cli {name}: {wall:.2f}s {phases}")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    rec = {
        "metric": "CLI wall seconds, 63k x 1142 TSV in -> full u16 TSV out", "value": runs["full_tsv"]["wall_s"],
        "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": runs["full_tsv"]["wall_s"] * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u16",
        "data": "synthetic (seeded long-tail Zipf(1.1) allele calls, Beta(0.5,15) per-sample missing)",
        "config": {"workload": "cli", "n": N, "L": L, "layout": "full", "stripe": 512, "tau": "auto"},
        "runs": runs,
    }
    print(json.dumps(rec), flush=True)




def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tau", type=int, default=0, help="headline tau (0 = cost model)")
    ap.add_argument("--tau-serve", type=int, default=0, help="serving database tau (0 = cost model)")
    ap.add_argument("--kernel", type=int, default=0,
                    help="CGH_KERNEL_*: 0 auto, 1 int8 MFMA, 2 fp4 MFMA, 3 VALU AND + popcount")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the batch and serve records")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration) and exit")
    ap.add_argument("--workload", default="all", choices=["all", "long_tail", "batch", "serve", "cli", "split"],
                    help="all: the headline line with its batch and serve records; long_tail: the headline "
                         "alone; batch / serve / cli / split: that record alone")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests in the stream")
    ap.add_argument("--summarize-counters", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="no GPU: turn four rocprofv3 --pmc csv directories into "
                         "profiles/hbm_traffic_long_tail_63005x1142.json (needs --tau)")
    args = ap.parse_args()
    if args.summarize_counters:
        return summarize_counters(args.summarize_counters, args.tau, 2 if args.kernel == 0 else args.kernel,
                                  os.path.join(ROOT, "profiles", "hbm_traffic_long_tail_63005x1142.json"))
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the distance path runs on a single device "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    import torch

    torch.zeros(1, device="cuda")  # torch's HIP runtime first; the library shares it
    if args.copy_calibration:
        src = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 1 << 30, "copies": 3}), flush=True)
        return
    cgh = load_package()
    import cgmlst_hamming_gpu_amd.workloads as wl
    if args.workload == "batch":
        print(json.dumps(batch_record(args, cgh, wl, torch)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(serve_record(args, cgh, wl, torch)), flush=True)
        return
    if args.workload == "cli":
        return run_cli(args, cgh, wl)
    if args.workload == "split":
        print(json.dumps(split_record(args, cgh, wl, torch)), flush=True)
        return

    rec, X, flat = headline(args, cgh, wl, torch)
    if args.workload == "all" and not args.headline_only:
        rec["batch"] = batch_record(args, cgh, wl, torch)
        rec["serve"] = serve_record(args, cgh, wl, torch)
    cpu = None if args.no_cpu_baseline else cpu_baseline(X, flat, cgh, wl)
    if cpu is not None:
        log(f"cpu baseline {cpu['value']:.4f} Gpairs/s on {cpu['cores']} threads ({cpu['cpu_model']}), "
            f"2,000 x 1,142 in full {cpu['config1_full']['gpairs_s']:.4f} Gpairs/s, "
            f"sample matches GPU: {cpu['sample_matches_gpu']}")
    rec["cpu_baseline"] = cpu
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
