"""Python binding of libcgmlst_hamming.so (ctypes over the C ABI in
include/cgmlst_hamming.h).

The product is the C ABI library and the `cgmlst-hamming-gpu` command line;
this module is a thin host-side mirror used by the tests and the benchmark.
It mirrors the reference's host pipeline (tensorcore/rs/src/main.rs:353-465:
dense ranks -> tau -> O/P bits + rare pairs -> GPU stripes) with the same
argument meaning and error behaviour (failures raise CghError carrying the
library's message, the reference's `die` text).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the distance computation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcgmlst_hamming.so")
CLI_PATH = os.path.join(HERE, "bin", "cgmlst-hamming-gpu")

# Symbols the header declares, in declaration order.
EXPORTS = (
    "tch_device_count", "tch_free_vram", "tch_run",
    "cgh_dense_ranks", "cgh_encode", "cgh_encode_meta", "cgh_column_table", "cgh_rare_csr", "cgh_encoding_free",
    "cgh_allele_counts", "cgh_choose_tau",
    "cgh_dataset_upload", "cgh_dataset_upload_split", "cgh_dataset_build", "cgh_dataset_encode", "cgh_dataset_set_cap", "cgh_dataset_autotune", "cgh_dataset_info", "cgh_dataset_free", "cgh_upper_elems", "cgh_stripe_offset",
    "cgh_dataset_upper", "cgh_dataset_phase", "cgh_dataset_stream", "cgh_dataset_stream_graph",
    "cgh_batch_create", "cgh_batch_out_elems", "cgh_batch_out_offset", "cgh_batch_tile", "cgh_batch_run",
    "cgh_batch_free",
    "cgh_codec_create", "cgh_codec_info", "cgh_codec_members", "cgh_codec_encode", "cgh_codec_encode_rows",
    "cgh_codec_free",
    "cgh_db_create", "cgh_db_info", "cgh_db_free",
    "cgh_server_create", "cgh_server_slots", "cgh_server_submit", "cgh_server_wait", "cgh_server_free",
    "cgh_measure_peaks", "cgh_measure_memory", "cgh_dataset_tile", "cgh_batch_run_noalloc",
    "cgh_dataset_tuning", "cgh_batch_autotune",
)

PHASE_PRODUCT, PHASE_RARE = 1, 2
WANT_DIST = 1


class CghError(RuntimeError):
    pass


class Encoding(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("L", C.c_int), ("tau", C.c_int),
        ("n_pad", C.c_size_t), ("D", C.c_size_t), ("kw1", C.c_size_t), ("kw2", C.c_size_t),
        ("O", C.POINTER(C.c_uint32)), ("P", C.POINTER(C.c_uint32)),
        ("n_groups", C.c_size_t), ("n_members", C.c_size_t),
        ("gstart", C.POINTER(C.c_uint32)), ("gmem", C.POINTER(C.c_uint32)),
        ("gpairs", C.POINTER(C.c_uint64)), ("n_rare", C.c_uint64),
        ("n_rentries", C.c_size_t), ("rstart", C.POINTER(C.c_uint32)), ("rentry", C.POINTER(C.c_uint64)),
    ]


TCH_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.POINTER(C.c_uint16))
LOG_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)

_lib = None


def load():
    """Load the built library (building is __graft_entry__.build()'s job)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CghError(f"{LIB_PATH} is not built: run __graft_entry__.build() (make -C cgmlst-hamming-gpu_amd)")
    lib = C.CDLL(LIB_PATH)
    P, S, I, U64 = C.c_void_p, C.c_size_t, C.c_int, C.c_uint64
    sig = {
        "tch_device_count": (I, []),
        "tch_free_vram": (S, [I]),
        "tch_run": (I, [I, I, S, P, S, P, S, P, P, S, I, TCH_CB, P, C.POINTER(C.c_double), C.POINTER(I), C.c_char_p, I]),
        "cgh_dense_ranks": (I, [P, I, I, P, C.c_char_p, I]),
        "cgh_encode": (I, [P, I, I, I, I, C.POINTER(Encoding), C.c_char_p, I]),
        "cgh_encode_meta": (I, [P, I, I, I, I, C.POINTER(Encoding), C.c_char_p, I]),
        "cgh_column_table": (S, [P, I, I, I, P, P, S]),
        "cgh_dataset_build": (I, [I, P, I, I, I, C.POINTER(P), C.c_char_p, I]),
        "cgh_dataset_encode": (I, [I, P, I, I, I, S, P, P, C.POINTER(P), C.POINTER(I), P, C.c_char_p, I]),
        "cgh_dataset_set_cap": (I, [P, I]),
        "cgh_dataset_autotune": (I, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_double), C.c_char_p, I]),
        "cgh_dataset_info": (None, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(S), C.POINTER(U64)]),
        "cgh_rare_csr": (I, [C.POINTER(Encoding), P, P, C.c_char_p, I]),
        "cgh_encoding_free": (None, [C.POINTER(Encoding)]),
        "cgh_allele_counts": (S, [P, I, I, P, S]),
        "cgh_choose_tau": (I, [P, S, I, I, S, S, LOG_CB, P]),
        "cgh_dataset_upload": (I, [I, C.POINTER(Encoding), C.POINTER(P), C.c_char_p, I]),
        "cgh_dataset_upload_split": (I, [I, C.POINTER(Encoding), S, C.POINTER(P), C.c_char_p, I]),
        "cgh_dataset_free": (None, [P]),
        "cgh_upper_elems": (S, [I, I]),
        "cgh_stripe_offset": (S, [I, I, I]),
        "cgh_dataset_upper": (I, [P, I, P, P, I, C.c_char_p, I]),
        "cgh_dataset_phase": (I, [P, I, P, P, I, I, C.c_char_p, I]),
        "cgh_dataset_stream": (I, [P, I, TCH_CB, P, C.POINTER(C.c_double), C.c_char_p, I]),
        "cgh_dataset_stream_graph": (I, [P, I, TCH_CB, P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p,
                                         I]),
        "cgh_measure_peaks": (I, [I] + [C.POINTER(C.c_double)] * 4 + [C.c_char_p, I]),
        "cgh_batch_create": (I, [I, C.POINTER(P), I, I, I, C.POINTER(P), C.c_char_p, I]),
        "cgh_batch_out_elems": (S, [P]),
        "cgh_batch_out_offset": (S, [P, I]),
        "cgh_batch_tile": (I, [P, I]),
        "cgh_batch_run": (I, [P, P, P, I, I, C.c_char_p, I]),
        "cgh_batch_free": (None, [P]),
        "cgh_codec_create": (I, [P, I, I, I, C.POINTER(Encoding), C.POINTER(P), C.c_char_p, I]),
        "cgh_codec_info": (None, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)] + [C.POINTER(S)] * 4),
        "cgh_codec_members": (C.POINTER(C.c_uint32), [P]),
        "cgh_codec_encode": (C.c_long, [P, P, I, P, P, P, P, S, C.c_char_p, I]),
        "cgh_codec_encode_rows": (C.c_long, [P, P, I, P, P, P, P, S, C.c_char_p, I]),
        "cgh_codec_free": (None, [P]),
        "cgh_db_create": (I, [I, P, I, I, I, C.POINTER(P), C.c_char_p, I]),
        "cgh_db_info": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(S)]),
        "cgh_db_free": (None, [P]),
        "cgh_server_create": (I, [P, I, I, S, I, C.POINTER(P), C.c_char_p, I]),
        "cgh_server_slots": (I, [P]),
        "cgh_server_submit": (C.c_long, [P, P, I, I, C.c_char_p, I]),
        "cgh_server_wait": (I, [P, C.c_long, P, P, P, C.POINTER(C.c_double), C.c_char_p, I]),
        "cgh_server_free": (None, [P]),
        "cgh_dataset_tile": (I, [P, I]),
        "cgh_batch_run_noalloc": (I, [C.POINTER(P), I, I, P, P, I, C.c_char_p, I]),
        "cgh_measure_memory": (I, [I] + [C.POINTER(C.c_double)] * 3 + [C.c_char_p, I]),
        "cgh_dataset_tuning": (None, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "cgh_batch_autotune": (I, [P, P, P, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_double), C.c_char_p, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _err():
    return C.create_string_buffer(512)


def _check(rc, err):
    if rc:
        raise CghError(err.value.decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    return load().tch_device_count()


def free_vram(device: int = 0) -> int:
    return load().tch_free_vram(device)


def measure_peaks(device: int = 0) -> dict:
    """Register-only peaks of int8 MFMA, fp4 MFMA, VALU AND+popcount and a
    streaming copy (cgh_measure_peaks)."""
    v = [C.c_double() for _ in range(4)]
    err = _err()
    _check(load().cgh_measure_peaks(device, *[C.byref(x) for x in v], err, 512), err)
    return {"int8_mfma_tops": v[0].value, "fp4_mfma_tflops": v[1].value, "valu_popcount_tops": v[2].value,
            "copy_gbs": v[3].value}


def measure_memory(device: int = 0) -> dict:
    """Streaming copy, read-only stream and cache-resident read rates in GB/s
    (cgh_measure_memory)."""
    v = [C.c_double() for _ in range(3)]
    err = _err()
    _check(load().cgh_measure_memory(device, *[C.byref(x) for x in v], err, 512), err)
    return {"copy_gbs": v[0].value, "read_gbs": v[1].value, "cache_read_gbs": v[2].value}


def dense_ranks(calls: np.ndarray) -> np.ndarray:
    """u64 allele ids (n x L, 0 = missing) -> u16 dense per-locus ranks
    (main.rs:135-165)."""
    calls = np.ascontiguousarray(calls, dtype=np.uint64)
    n, L = calls.shape
    X = np.zeros((n, L), dtype=np.uint16)
    err = _err()
    _check(load().cgh_dense_ranks(_ptr(calls), n, L, _ptr(X), err, 512), err)
    return X


def allele_counts(X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.uint16)
    n, L = X.shape
    lib = load()
    need = lib.cgh_allele_counts(_ptr(X), n, L, None, 0)
    cnt = np.zeros(max(need, 1), dtype=np.uint32)
    lib.cgh_allele_counts(_ptr(X), n, L, _ptr(cnt), need)
    return cnt[:need]


def choose_tau(X: np.ndarray, out_bytes: int, free_bytes: int, log=None) -> int:
    """Auto tau (main.rs:169-198 rule, MI355X rates) for a run that
    allocates out_bytes of output buffers; 0 if none fits."""
    n, L = X.shape
    cnt = allele_counts(X)
    lines = []
    cb = LOG_CB(lambda ctx, line: lines.append(line.decode()))
    tau = load().cgh_choose_tau(_ptr(cnt), cnt.size, n, L, out_bytes, free_bytes, cb, None)
    if log is not None:
        for ln in lines:
            log(ln)
    return tau


class Encoded:
    """Host operands of one profile set (cgh_encode)."""

    def __init__(self, X: np.ndarray, tau: int, k_align: int = 256):
        X = np.ascontiguousarray(X, dtype=np.uint16)
        n, L = X.shape
        self.enc = Encoding()
        err = _err()
        _check(load().cgh_encode(_ptr(X), n, L, int(tau), int(k_align), C.byref(self.enc), err, 512), err)

    def __del__(self):
        if _lib is not None and getattr(self, "enc", None) is not None:
            _lib.cgh_encoding_free(C.byref(self.enc))
            self.enc = None

    def _arr(self, field, count, dtype):
        p = getattr(self.enc, field)
        if not p or count == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(p, shape=(count,)).copy()

    @property
    def O(self):
        e = self.enc
        return self._arr("O", e.n_pad * e.kw1, np.uint32).reshape(e.n_pad, e.kw1)

    @property
    def P(self):
        e = self.enc
        return self._arr("P", e.n_pad * e.kw2, np.uint32).reshape(e.n_pad, e.kw2)

    def groups(self):
        e = self.enc
        return (self._arr("gstart", e.n_groups + 1, np.uint32), self._arr("gmem", e.n_members, np.uint32),
                self._arr("gpairs", e.n_groups + 1, np.uint64))

    def row_index(self):
        """(rstart[n+1], rentry) -> list of partner arrays per row."""
        e = self.enc
        rstart = self._arr("rstart", e.n + 1, np.uint32)
        rentry = self._arr("rentry", e.n_rentries, np.uint64)
        gmem = self._arr("gmem", e.n_members, np.uint32)
        out = []
        for i in range(e.n):
            parts = [gmem[int(r & 0xFFFFFFFF):int(r >> 32)] for r in rentry[rstart[i]:rstart[i + 1]]]
            out.append(np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint32))
        return out

    def rare_csr(self):
        e = self.enc
        start = np.zeros(e.n + 1, dtype=np.uint32)
        js = np.zeros(max(e.n_rare, 1), dtype=np.uint32)
        err = _err()
        _check(load().cgh_rare_csr(C.byref(self.enc), _ptr(start), _ptr(js), err, 512), err)
        return start, js[: e.n_rare]


def upper_elems(n: int, S: int) -> int:
    return load().cgh_upper_elems(n, S)


def stripe_offset(n: int, S: int, s: int) -> int:
    return load().cgh_stripe_offset(n, S, s)


class Dataset:
    """Device-resident operands (cgh_dataset_upload)."""

    def __init__(self, enc: Encoded = None, device: int = 0, o_bytes: int = 0):
        """o_bytes > 0: cgh_dataset_upload_split (the one-hot stays in host
        memory and is streamed through o_bytes of device buffers)."""
        self.handle = C.c_void_p()
        if enc is None:
            return
        self.n = enc.enc.n
        err = _err()
        if o_bytes:
            _check(load().cgh_dataset_upload_split(device, C.byref(enc.enc), o_bytes, C.byref(self.handle), err, 512),
                   err)
        else:
            _check(load().cgh_dataset_upload(device, C.byref(enc.enc), C.byref(self.handle), err, 512), err)

    @classmethod
    def build(cls, X: np.ndarray, tau: int, device: int = 0):
        """cgh_dataset_build: O/P built on the device from dense ranks X."""
        X = np.ascontiguousarray(X, dtype=np.uint16)
        ds = cls()
        ds.n = X.shape[0]
        err = _err()
        _check(load().cgh_dataset_build(device, _ptr(X), X.shape[0], X.shape[1], int(tau), C.byref(ds.handle),
                                        err, 512), err)
        return ds

    @classmethod
    def encode(cls, calls: np.ndarray, tau: int = 0, out_bytes: int = 0, device: int = 0, want_ranks: bool = False):
        """cgh_dataset_encode: u64 ids encoded on the device.  -> (dataset,
        tau used, dense ranks or None)"""
        calls = np.ascontiguousarray(calls, dtype=np.uint64)
        n, L = calls.shape
        X = np.zeros((n, L), dtype=np.uint16) if want_ranks else None
        ds = cls()
        ds.n = n
        t = C.c_int()
        err = _err()
        _check(load().cgh_dataset_encode(device, _ptr(calls), n, L, int(tau), int(out_bytes), None, None,
                                         C.byref(ds.handle), C.byref(t), _ptr(X) if X is not None else None, err,
                                         512), err)
        return ds, t.value, X

    def autotune(self, S: int):
        """cgh_dataset_autotune -> (kernel, tile, ms of the fastest)"""
        k, t, ms = C.c_int(), C.c_int(), C.c_double()
        err = _err()
        _check(load().cgh_dataset_autotune(self.handle, S, C.byref(k), C.byref(t), C.byref(ms), err, 512), err)
        return k.value, t.value, ms.value

    def tuning(self):
        """cgh_dataset_tuning -> (kernel, tile, LDS ring depth) that CGH_KERNEL_AUTO runs."""
        k, t, st = C.c_int(), C.c_int(), C.c_int()
        load().cgh_dataset_tuning(self.handle, C.byref(k), C.byref(t), C.byref(st))
        return k.value, t.value, st.value

    def set_cap(self, cap: int):
        """Outputs from now on: values above cap written as cap (0 = none)."""
        if load().cgh_dataset_set_cap(self.handle, int(cap)):
            raise CghError(f"bad cap {cap}")

    def info(self):
        """-> dict(n, L, tau, D, n_rare)"""
        v = [C.c_int(), C.c_int(), C.c_int(), C.c_size_t(), C.c_uint64()]
        load().cgh_dataset_info(self.handle, *[C.byref(x) for x in v])
        return dict(zip(("n", "L", "tau", "D", "n_rare"), (x.value for x in v)))

    def close(self):
        if _lib is not None and self.handle:
            _lib.cgh_dataset_free(self.handle)
            self.handle = C.c_void_p()

    __del__ = close

    def stream(self, S: int, graph: bool = False):
        """cgh_dataset_stream (graph: cgh_dataset_stream_graph) into a square
        symmetric H (host)."""
        n = self.n
        H = np.zeros((n, n), dtype=np.uint16)
        seen = []

        def cb(ctx, r0, rows, w, stripe):
            H[r0:r0 + rows, r0:] = np.ctypeslib.as_array(stripe, shape=(rows * w,)).reshape(rows, w)
            seen.append((r0, rows, w))

        ccb = TCH_CB(cb)
        secs = C.c_double()
        err = _err()
        if graph:
            build = C.c_double()
            _check(load().cgh_dataset_stream_graph(self.handle, S, ccb, None, C.byref(secs), C.byref(build), err, 512),
                   err)
        else:
            _check(load().cgh_dataset_stream(self.handle, S, ccb, None, C.byref(secs), err, 512), err)
        iu = np.triu_indices(n, 1)
        H.T[iu] = H[iu]
        np.fill_diagonal(H, 0)
        return H, secs.value, seen

    def tile(self, kernel: int = 0) -> int:
        return load().cgh_dataset_tile(self.handle, kernel)

    def upper(self, S: int, out_ptr: int, stream_ptr: int = 0, kernel: int = 0):
        """Enqueue the exact upper triangle into device memory out_ptr."""
        err = _err()
        _check(load().cgh_dataset_upper(self.handle, S, C.c_void_p(out_ptr), C.c_void_p(stream_ptr), kernel, err,
                                        512), err)

    def phase(self, S: int, out_ptr: int, phases: int, stream_ptr: int = 0, kernel: int = 0):
        """Enqueue PHASE_PRODUCT and/or PHASE_RARE (cgh_dataset_phase)."""
        err = _err()
        _check(load().cgh_dataset_phase(self.handle, S, C.c_void_p(out_ptr), C.c_void_p(stream_ptr), kernel, phases,
                                        err, 512), err)


class Batch:
    """Many uploaded datasets in one persistent launch (cgh_batch_*)."""

    def __init__(self, datasets, S: int, tile: int = 0, device: int = 0):
        self.datasets = list(datasets)  # keep them alive
        arr = (C.c_void_p * len(self.datasets))(*[d.handle for d in self.datasets])
        self.handle = C.c_void_p()
        err = _err()
        _check(load().cgh_batch_create(device, arr, len(self.datasets), S, tile, C.byref(self.handle), err, 512), err)
        self.S = S

    def close(self):
        if _lib is not None and self.handle:
            _lib.cgh_batch_free(self.handle)
            self.handle = C.c_void_p()

    __del__ = close

    def out_elems(self) -> int:
        return load().cgh_batch_out_elems(self.handle)

    def offset(self, d: int) -> int:
        return load().cgh_batch_out_offset(self.handle, d)

    def tile(self, d: int) -> int:
        return load().cgh_batch_tile(self.handle, d)

    def run(self, out_ptr: int, stream_ptr: int = 0, kernel: int = 0, graph: bool = False):
        err = _err()
        _check(load().cgh_batch_run(self.handle, C.c_void_p(out_ptr), C.c_void_p(stream_ptr), kernel, int(graph), err,
                                    512), err)

    def autotune(self, out_ptr: int, stream_ptr: int = 0):
        """cgh_batch_autotune -> (row threshold for 256-tiles, kernel, ms of the fastest)."""
        th, k, ms = C.c_int(), C.c_int(), C.c_double()
        err = _err()
        _check(load().cgh_batch_autotune(self.handle, C.c_void_p(out_ptr), C.c_void_p(stream_ptr), C.byref(th),
                                         C.byref(k), C.byref(ms), err, 512), err)
        return th.value, k.value, ms.value


def batch_run_noalloc(datasets, S: int, out_ptr: int, stream_ptr: int = 0, kernel: int = 0):
    """cgh_batch_run_noalloc: the whole batch of `datasets` in one call that
    allocates no device memory (the layout of Batch(datasets, S))."""
    arr = (C.c_void_p * len(datasets))(*[d.handle for d in datasets])
    err = _err()
    _check(load().cgh_batch_run_noalloc(arr, len(datasets), S, C.c_void_p(out_ptr), C.c_void_p(stream_ptr), kernel,
                                        err, 512), err)


class QueryCodec:
    """Host codec of a database (cgh_codec_*): query ids -> the database's
    frequent columns and rare member ranges.  Host memory only."""

    def __init__(self, calls: np.ndarray, tau: int, want_encoding: bool = False):
        calls = np.ascontiguousarray(calls, dtype=np.uint64)
        n, L = calls.shape
        self.handle = C.c_void_p()
        self.enc = Encoding() if want_encoding else None
        err = _err()
        _check(load().cgh_codec_create(_ptr(calls), n, L, int(tau), C.byref(self.enc) if want_encoding else None,
                                       C.byref(self.handle), err, 512), err)
        v = [C.c_int() for _ in range(3)] + [C.c_size_t() for _ in range(4)]
        load().cgh_codec_info(self.handle, *[C.byref(x) for x in v])
        self.n, self.L, self.tau, self.kw1, self.kw2, self.D, self.n_members = [x.value for x in v]

    def members(self) -> np.ndarray:
        if not self.n_members:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(load().cgh_codec_members(self.handle), shape=(self.n_members,)).copy()

    def encode(self, calls: np.ndarray):
        """-> (O q x kw1 u32, P q x kw2 u32, rstart q+1 u32, rentry u64)"""
        calls = np.ascontiguousarray(calls, dtype=np.uint64)
        q = calls.shape[0]
        O = np.full((q, self.kw1), 0xFFFFFFFF, dtype=np.uint32)  # overwritten rows
        P = np.full((q, self.kw2), 0xFFFFFFFF, dtype=np.uint32)
        rs = np.zeros(q + 1, dtype=np.uint32)
        re = np.zeros(max(q * self.L, 1), dtype=np.uint64)
        err = _err()
        m = load().cgh_codec_encode(self.handle, _ptr(calls), q, _ptr(O), _ptr(P), _ptr(rs), _ptr(re), re.size,
                                    err, 512)
        if m < 0:
            raise CghError(err.value.decode(errors="replace"))
        return O, P, rs, re[:m].copy()

    def encode_rows(self, parts):
        """encode() of the queries of several arrays (each q_i x L) in one
        call through cgh_codec_encode_rows, as if they were stacked."""
        parts = [np.ascontiguousarray(a, dtype=np.uint64) for a in parts]
        rows = [a[i].ctypes.data for a in parts for i in range(a.shape[0])]
        q = len(rows)
        ptrs = (C.c_void_p * max(q, 1))(*rows)
        O = np.full((q, self.kw1), 0xFFFFFFFF, dtype=np.uint32)
        P = np.full((q, self.kw2), 0xFFFFFFFF, dtype=np.uint32)
        rs = np.zeros(q + 1, dtype=np.uint32)
        re = np.zeros(max(q * self.L, 1), dtype=np.uint64)
        err = _err()
        m = load().cgh_codec_encode_rows(self.handle, C.cast(ptrs, C.c_void_p), q, _ptr(O), _ptr(P), _ptr(rs),
                                         _ptr(re), re.size, err, 512)
        if m < 0:
            raise CghError(err.value.decode(errors="replace"))
        return O, P, rs, re[:m].copy()

    def close(self):
        if _lib is not None and self.handle:
            _lib.cgh_codec_free(self.handle)
            self.handle = C.c_void_p()
        if _lib is not None and self.enc is not None:
            _lib.cgh_encoding_free(C.byref(self.enc))
            self.enc = None

    __del__ = close


class Database:
    """Resident database (cgh_db_*)."""

    def __init__(self, calls: np.ndarray, tau: int, device: int = 0):
        calls = np.ascontiguousarray(calls, dtype=np.uint64)
        n, L = calls.shape
        self.handle = C.c_void_p()
        err = _err()
        _check(load().cgh_db_create(device, _ptr(calls), n, L, int(tau), C.byref(self.handle), err, 512), err)
        v = [C.c_int(), C.c_int(), C.c_int()]
        d = C.c_size_t()
        load().cgh_db_info(self.handle, *[C.byref(x) for x in v], C.byref(d))
        self.n, self.L, self.tau, self.D = v[0].value, v[1].value, v[2].value, d.value

    def close(self):
        if _lib is not None and self.handle:
            _lib.cgh_db_free(self.handle)
            self.handle = C.c_void_p()

    __del__ = close


class Server:
    """Continuous-batching query server over a Database (cgh_server_*)."""

    def __init__(self, db: Database, slots: int = 4, max_batch: int = 1024, hbm_budget: int = 0, kernel: int = 0):
        self.db = db
        self.handle = C.c_void_p()
        err = _err()
        _check(load().cgh_server_create(db.handle, slots, max_batch, hbm_budget, kernel, C.byref(self.handle), err,
                                        512), err)
        self.slots = load().cgh_server_slots(self.handle)
        self._tickets = {}

    def close(self):
        if _lib is not None and self.handle:
            _lib.cgh_server_free(self.handle)
            self.handle = C.c_void_p()

    __del__ = close

    def submit(self, calls: np.ndarray, want_dist: bool = True) -> int:
        calls = np.ascontiguousarray(calls, dtype=np.uint64)
        err = _err()
        t = load().cgh_server_submit(self.handle, _ptr(calls), calls.shape[0], WANT_DIST if want_dist else 0, err,
                                     512)
        if t < 0:
            raise CghError(err.value.decode(errors="replace"))
        self._tickets[t] = (calls.shape[0], want_dist)
        return t

    def wait(self, ticket: int):
        """-> (dist q x n u16, or None when submitted without it; top_d q x
        10; top_i q x 10; latency s)"""
        q, want_dist = self._tickets.pop(ticket)
        dist = np.zeros((q, self.db.n), dtype=np.uint16) if want_dist else None
        top_d = np.zeros((q, 10), dtype=np.uint16)
        top_i = np.zeros((q, 10), dtype=np.uint32)
        lat = C.c_double()
        err = _err()
        _check(load().cgh_server_wait(self.handle, ticket, _ptr(dist) if dist is not None else None, _ptr(top_d),
                                      _ptr(top_i), C.byref(lat), err, 512), err)
        return dist, top_d, top_i, lat.value


def upper_to_square(flat: np.ndarray, n: int, S: int) -> np.ndarray:
    """Stripe-layout upper triangle (u16) -> symmetric n x n matrix."""
    H = np.zeros((n, n), dtype=np.uint16)
    for s, r0 in enumerate(range(0, n, S)):
        rows, w = min(S, n - r0), n - r0
        off = stripe_offset(n, S, s)
        blk = flat[off: off + rows * w].reshape(rows, w)
        H[r0:r0 + rows, r0:] = blk
    iu = np.triu_indices(n, 1)
    H.T[iu] = H[iu]
    np.fill_diagonal(H, 0)
    return H


def tch_run(enc: Encoded, S: int = 512, device: int = 0, keep_rows=None, shuffle_seed=None):
    """Run the reference-compatible tch_run on the encoding (K aligned to
    1024).  Returns (H, gpu_secs, rare_on_gpu, stripes seen): H is the square
    symmetric matrix, or with keep_rows a dict row -> its cells j > row.
    shuffle_seed permutes each row's partner list first (the reference's CSR
    order within a row is arbitrary)."""
    e = enc.enc
    if (e.kw1 * 32) % 1024 or (e.kw2 * 32) % 1024:
        raise CghError("tch_run needs an encoding with k_align=1024")
    n = e.n
    start, js = enc.rare_csr()
    if shuffle_seed is not None:
        rng = np.random.default_rng(shuffle_seed)
        for i in range(n):
            rng.shuffle(js[start[i]:start[i + 1]])
    js_buf = js if js.size else np.zeros(1, dtype=np.uint32)
    H = np.zeros((n, n), dtype=np.uint16) if keep_rows is None else {}
    want = None if keep_rows is None else set(int(i) for i in keep_rows)
    seen = []

    def cb(ctx, r0, rows, w, stripe):
        blk = np.ctypeslib.as_array(stripe, shape=(rows * w,)).reshape(rows, w)
        if want is None:
            H[r0:r0 + rows, r0:] = blk
        else:
            for i in want:
                if r0 <= i < r0 + rows:
                    H[i] = blk[i - r0, i - r0 + 1:].copy()
        seen.append((r0, rows, w))

    ccb = TCH_CB(cb)
    secs = C.c_double()
    rare_gpu = C.c_int()
    err = _err()
    rc = load().tch_run(device, n, e.n_pad, C.cast(e.O, C.c_void_p), e.kw1 * 32, C.cast(e.P, C.c_void_p), e.kw2 * 32,
                        _ptr(start), _ptr(js_buf), int(e.n_rare), S, ccb, None, C.byref(secs), C.byref(rare_gpu),
                        err, 512)
    _check(rc, err)
    if want is None:
        iu = np.triu_indices(n, 1)
        H.T[iu] = H[iu]
        np.fill_diagonal(H, 0)
    return H, secs.value, bool(rare_gpu.value), seen
