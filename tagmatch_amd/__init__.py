"""Python bindings of the MI355X TagMatch engine (ctypes over the C ABI).

The engine is the native library tagmatch_amd/lib/libtagmatch_amd.so
(HIP kernels for gfx950 + C++ host code); this module only loads it and
moves numpy arrays across the C ABI declared in include/tagmatch_amd.h.
There is no Python or CPU fallback: every call fails loudly if the library
or the GPU is missing.
"""
import ctypes as C
import time
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtagmatch_amd.so")

TM_OK = 0
TM_ERR_INVALID = -1
TM_ERR_HIP = -2
TM_ERR_NOMEM = -3
TM_ERR_OVERFLOW = -4
TM_ERR_STATE = -5
TM_ERR_NODEV = -6
TM_ERR_BUSY = -7

MODE_KEYS_UNIQUE = 0
MODE_KEYS_ALL = 1
MODE_HITS = 2
MODE_KEY_BITMAP = 3
STATIC_MAX_QUERIES = 65536
STATIC_MAX_PARTS = 16384
SIZE_BUCKETS = 6
ROW_FULL = 0xFFFFFFFF
SCHED_PLAIN = 0
SCHED_PERSISTENT = 1


IO_BINARY, IO_ASCII, IO_BINARY_PARTITIONED, IO_ASCII_PARTITIONED = 0, 1, 2, 3
KIND_FIB, KIND_QUERIES, KIND_PREFIXES = 0, 1, 2


class TagMatchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class WorkloadParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n_tags", C.c_uint32), ("zipf_s", C.c_double),
        ("bits_per_tag", C.c_uint32), ("ftag_min", C.c_uint32), ("ftag_max", C.c_uint32),
        ("key_min", C.c_uint32), ("key_max", C.c_uint32), ("trees", C.c_uint32),
        ("ifaces", C.c_uint32), ("qtag_min", C.c_uint32), ("qtag_max", C.c_uint32),
        ("superset_frac", C.c_double),
    ]


class MatchConfig(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32), ("schedule", C.c_uint32), ("batch_size", C.c_uint32),
        ("chunk_blocks", C.c_uint32), ("waves_per_group", C.c_uint32),
        ("groups_per_cu", C.c_uint32), ("hit_capacity", C.c_uint64), ("use_graph", C.c_uint32),
        ("log_units", C.c_uint32), ("order_by_cost", C.c_uint32), ("variant", C.c_uint32),
        ("full_exit", C.c_uint32), ("compact_tiles", C.c_uint32), ("reserved", C.c_uint32 * 4),
    ]


class Status(C.Structure):
    _fields_ = [
        ("n_queries", C.c_uint32), ("n_items", C.c_uint32), ("n_pairs", C.c_uint64),
        ("n_hits", C.c_uint64), ("n_results", C.c_uint64), ("hit_capacity", C.c_uint64),
        ("overflow", C.c_int32), ("error", C.c_int32), ("device_ms", C.c_double),
        ("match_ms", C.c_double), ("n_block_visits", C.c_uint64), ("n_cand_tests", C.c_uint64),
        ("n_hit_events", C.c_uint64), ("pred_hits", C.c_double), ("model_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FibStats(C.Structure):
    _fields_ = [
        ("n_filters", C.c_uint64), ("n_blocks", C.c_uint64), ("n_parts", C.c_uint64),
        ("n_keys", C.c_uint64), ("n_distinct_keys", C.c_uint64), ("device_bytes", C.c_uint64),
        ("max_part_size", C.c_uint64),
    ]


class Prefix(C.Structure):
    _fields_ = [("words", C.c_uint32 * 6), ("length", C.c_uint32), ("partition", C.c_uint32), ("size", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("n_slots", C.c_uint32), ("slot_queries", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("flush_queries", C.c_uint32), ("flush_us", C.c_uint32), ("step_graphs", C.c_uint32),
                ("reserved", C.c_uint32 * 5)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "n_requests", "n_requests_done", "n_queries_done", "n_batches", "n_steps", "n_retries", "n_inserts",
        "n_live_inserts",
        "n_budget_waits", "requests_per_batch_sum", "reserved_peak", "reserved_now", "hbm_budget", "n_waiting")] + [
        ("latency_sum_ms", C.c_double), ("latency_max_ms", C.c_double), ("device_ms_sum", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_queries", C.c_uint32), ("n_values", C.c_uint64),
                ("step_queries", C.c_uint32), ("pad", C.c_uint32), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double)]


P = C.c_void_p
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)

_SIGS = {
    "tm_last_error": (C.c_char_p, []),
    "tm_version": (C.c_char_p, []),
    "tm_dev_init": (C.c_int, [C.c_int]),
    "tm_dev_shutdown": (C.c_int, []),
    "tm_dev_count": (C.c_int, [C.POINTER(C.c_int)]),
    "tm_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "tm_dev_synchronize": (C.c_int, []),
    "tm_dev_arch": (C.c_int, [C.c_char_p, C.c_size_t]),
    "tm_partition": (C.c_int, [P, C.c_uint32, C.c_uint32, P, P, P, C.c_uint32, U32P]),
    "tm_fib_create": (C.c_int, [C.POINTER(P)]),
    "tm_fib_destroy": (C.c_int, [P]),
    "tm_fib_upload": (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32, P, P, P]),
    "tm_fib_stats_get": (C.c_int, [P, C.POINTER(FibStats)]),
    "tm_fib_set_reserve": (C.c_int, [P, C.c_uint32]),
    "tm_fib_insert": (C.c_int, [P, P, C.c_uint32, P, P, P, C.c_uint32]),
    "tm_fib_insert_live": (C.c_int, [P, P, C.c_uint32, P, P, P]),
    "tm_match_config_default": (None, [C.POINTER(MatchConfig)]),
    "tm_ring_create": (C.c_int, [P, C.c_uint32, C.POINTER(MatchConfig), C.POINTER(P)]),
    "tm_ring_destroy": (C.c_int, [P]),
    "tm_ring_config": (C.c_int, [P, C.POINTER(MatchConfig)]),
    "tm_ring_query_buffer": (C.c_int, [P, C.POINTER(P)]),
    "tm_submit": (C.c_int, [P, P, C.c_uint32]),
    "tm_submit_resident": (C.c_int, [P, C.c_uint32]),
    "tm_submit_units": (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32]),
    "tm_poll": (C.c_int, [P]),
    "tm_wait": (C.c_int, [P, C.POINTER(Status)]),
    "tm_read_results": (C.c_int, [P, P, P, C.c_uint64]),
    "tm_ring_results_device": (C.c_int, [P, C.POINTER(P), C.POINTER(P)]),
    "tm_read_key_bitmaps": (C.c_int, [P, P, C.c_uint64, U32P]),
    "tm_ring_set_host_rows": (C.c_int, [P, P, C.c_uint64]),
    "tm_ring_set_host_rows_packed": (C.c_int, [P, P, C.c_uint64, P, C.c_uint64]),
    "tm_match_rows_static": (C.c_int, [P, P, C.c_uint32, P, P]),
    "tm_host_alloc": (C.c_int, [C.POINTER(P), C.c_uint64]),
    "tm_dev_alloc": (C.c_int, [C.POINTER(P), C.c_uint64]),
    "tm_dev_free": (C.c_int, [P]),
    "tm_dev_upload": (C.c_int, [P, P, C.c_uint64]),
    "tm_dev_download": (C.c_int, [P, P, C.c_uint64]),
    "tm_host_free": (C.c_int, [P]),
    "tm_fib_key_dictionary": (C.c_int, [P, P, C.c_uint64, U64P]),
    "tm_expand_key_bitmaps": (C.c_int, [P, C.c_uint64, C.c_uint32, P, C.c_uint32, P, P, C.c_uint64, C.c_uint32]),
    "tm_expand_key_rows_packed": (C.c_int, [P, P, C.c_uint64, C.c_uint64, C.c_uint32, P, C.c_uint32, P, P,
                                            C.c_uint64, C.c_uint32]),
    "tm_ring_lookup_counts": (C.c_int, [P, P, C.c_uint32]),
    "tm_match": (C.c_int, [P, P, C.c_uint32, C.POINTER(MatchConfig), P, P, C.c_uint64, C.POINTER(Status)]),
    "tm_mlp_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(P)]),
    "tm_mlp_destroy": (C.c_int, [P]),
    "tm_mlp_param_count": (C.c_uint64, [P]),
    "tm_mlp_get_params": (C.c_int, [P, P, C.c_uint64]),
    "tm_mlp_set_params": (C.c_int, [P, P, C.c_uint64]),
    "tm_mlp_forward": (C.c_int, [P, P, C.c_uint32, P]),
    "tm_mlp_loss_grad": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(C.c_float), P]),
    "tm_mlp_train": (C.c_int, [P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                               C.POINTER(C.c_float)]),
    "tm_ring_autotune": (C.c_int, [P, C.c_uint32, C.POINTER(MatchConfig), C.c_uint32, C.c_uint32, U32P, P]),
    "tm_ring_unit_log": (C.c_int, [P, P, P, P, P, P, C.c_uint32, U32P]),
    "tm_ring_set_bucket_segments": (C.c_int, [P, P, C.c_uint32]),
    "tm_ring_get_bucket_segments": (C.c_int, [P, P, C.c_uint32]),
    "tm_ring_autotune_buckets": (C.c_int, [P, P, P, C.c_uint32, P, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]),
    "tm_ring_set_cost_model": (C.c_int, [P, P, C.c_uint32]),
    "tm_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "tm_server_create": (C.c_int, [P, C.POINTER(ServerConfig), C.POINTER(MatchConfig), C.POINTER(P)]),
    "tm_server_destroy": (C.c_int, [P]),
    "tm_server_set_cost_model": (C.c_int, [P, P, C.c_uint32]),
    "tm_server_submit": (C.c_int, [P, P, C.c_uint32, U64P]),
    "tm_server_insert": (C.c_int, [P, P, C.c_uint32, P, P, P, C.c_uint32, U64P]),
    "tm_server_wait": (C.c_int, [P, C.c_uint64, C.c_double]),
    "tm_server_collect": (C.c_int, [P, P, C.c_uint32, U32P]),
    "tm_server_result": (C.c_int, [P, C.c_uint64, C.POINTER(RequestInfo), P, P, C.c_uint64]),
    "tm_server_release": (C.c_int, [P, C.c_uint64]),
    "tm_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
    "tm_canonicalize": (C.c_int, [P, C.c_uint64, P, P, C.POINTER(P), U64P, C.POINTER(P), C.POINTER(P)]),
    "tm_free": (None, [P]),
    "tm_workload_default_params": (None, [C.POINTER(WorkloadParams)]),
    "tm_gen_create": (P, [C.POINTER(WorkloadParams)]),
    "tm_gen_destroy": (None, [P]),
    "tm_gen_tag_bits": (C.c_int, [P, C.c_uint32, P]),
    "tm_gen_filters": (C.c_int, [P, C.c_uint64, C.c_uint64, P, P, P, C.c_int]),
    "tm_gen_queries": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_uint64, P, P, C.c_int]),
    "tm_fib_save": (C.c_int, [P, C.c_char_p]),
    "tm_fib_load": (C.c_int, [P, C.c_char_p]),
    "tm_io_scan_fib": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "tm_io_read_fib": (C.c_int, [C.c_char_p, C.c_uint32, P, P, P, P, C.c_uint64, C.c_uint64]),
    "tm_io_write_fib": (C.c_int, [C.c_char_p, C.c_uint32, P, C.c_uint64, P, P, P]),
    "tm_io_scan_queries": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "tm_io_read_queries": (C.c_int, [C.c_char_p, C.c_uint32, P, C.c_uint64]),
    "tm_io_write_queries": (C.c_int, [C.c_char_p, C.c_uint32, P, C.c_uint64]),
    "tm_io_scan_prefixes": (C.c_int, [C.c_char_p, C.c_uint32, U64P]),
    "tm_io_read_prefixes": (C.c_int, [C.c_char_p, C.c_uint32, P, C.c_uint64]),
    "tm_io_write_prefixes": (C.c_int, [C.c_char_p, C.c_uint32, P, C.c_uint64]),
    "tm_io_convert": (C.c_int, [C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_uint32, U64P]),
    "tm_io_sample": (C.c_int, [C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32,
                               U64P]),
    "tm_io_split": (C.c_int, [C.c_char_p, C.c_uint32, C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32, U64P]),
}

# Declared in include/tagmatch_amd.h; tests check every one is exported.
EXPORTED = sorted(_SIGS)

_lib = None


def load():
    """Load libtagmatch_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TagMatchError(TM_ERR_NODEV, f"{LIB_PATH} is missing: run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


def check(rc, what=""):
    if rc != TM_OK:
        msg = load().tm_last_error().decode(errors="replace")
        raise TagMatchError(rc, f"{what}: {msg}" if what else msg)
    return rc


def _threads():
    return int(os.environ.get("TM_HOST_THREADS", "0")) or min(16, os.cpu_count() or 1)


# ------------------------------------------------------------------ device

def init(device=0):
    check(load().tm_dev_init(device), "tm_dev_init")


def device_count():
    n = C.c_int(0)
    rc = load().tm_dev_count(C.byref(n))
    return n.value if rc == TM_OK else 0


def arch():
    buf = C.create_string_buffer(64)
    check(load().tm_dev_arch(buf, 64), "tm_dev_arch")
    return buf.value.decode()


def mem_info():
    f, t = C.c_size_t(0), C.c_size_t(0)
    check(load().tm_mem_info(C.byref(f), C.byref(t)), "tm_mem_info")
    return f.value, t.value


# ------------------------------------------------------------------ workload

def default_params(**kw):
    p = WorkloadParams()
    load().tm_workload_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Generator:
    """Seeded workload generator (tm_gen_*)."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_params(**kw)
        self.h = load().tm_gen_create(C.byref(self.params))
        if not self.h:
            raise TagMatchError(TM_ERR_INVALID, "tm_gen_create rejected the parameters")

    def __del__(self):
        if getattr(self, "h", None):
            load().tm_gen_destroy(self.h)
            self.h = None

    def filters(self, n, first=0):
        """Raw generated filters: words (n,6), key_off (n+1,), keys."""
        kmax = self.params.key_max
        w = np.zeros((n, 6), np.uint32)
        kc = np.zeros(n, np.uint32)
        k = np.zeros((n, kmax), np.uint32)
        check(load().tm_gen_filters(self.h, first, n, _ptr(w), _ptr(kc), _ptr(k), _threads()), "tm_gen_filters")
        key_off = np.zeros(n + 1, np.uint64)
        np.cumsum(kc, out=key_off[1:])
        mask = np.arange(kmax)[None, :] < kc[:, None]
        return w, key_off, k[mask].astype(np.uint32)

    def queries(self, n, n_filters, first=0):
        q = np.zeros((n, 6), np.uint32)
        s = np.zeros(n, np.uint8)
        check(load().tm_gen_queries(self.h, n_filters, first, n, _ptr(q), _ptr(s), _threads()), "tm_gen_queries")
        return q, s


def canonicalize(words, key_off, keys):
    """Merge identical filters, sort lexicographically (tm_canonicalize)."""
    words = np.ascontiguousarray(words, np.uint32)
    key_off = np.ascontiguousarray(key_off, np.uint64)
    keys = np.ascontiguousarray(keys, np.uint32)
    n = words.shape[0]
    ow, ok, ov = C.c_void_p(), C.c_void_p(), C.c_void_p()
    on = C.c_uint64()
    lib = load()
    check(lib.tm_canonicalize(_ptr(words), n, _ptr(key_off), _ptr(keys), C.byref(ow), C.byref(on),
                              C.byref(ok), C.byref(ov)), "tm_canonicalize")
    m = on.value
    try:
        cw = np.ctypeslib.as_array(C.cast(ow, U32P), shape=(max(m, 1) * 6,))[: m * 6].copy().reshape(m, 6)
        ck = np.ctypeslib.as_array(C.cast(ok, U64P), shape=(m + 1,)).copy()
        nk = int(ck[-1])
        cv = np.ctypeslib.as_array(C.cast(ov, U32P), shape=(max(nk, 1),))[:nk].copy()
    finally:
        lib.tm_free(ow)
        lib.tm_free(ok)
        lib.tm_free(ov)
    return cw, ck, cv


def workload(n_filters, n_queries, seed=None, query_first=0, **kw):
    """Generate, canonicalize: (words, key_off, keys, queries, superset)."""
    if seed is not None:
        kw["seed"] = seed
    g = Generator(**kw)
    w, ko, k = g.filters(n_filters)
    q, s = g.queries(n_queries, n_filters, first=query_first)
    cw, ck, cv = canonicalize(w, ko, k)
    return cw, ck, cv, q, s


# ------------------------------------------------------------------ partition / FIB

def partition(words, maxp, max_parts=None):
    """GPU balanced partitioning: (perm, part_offsets, masks)."""
    words = np.ascontiguousarray(words, np.uint32)
    n = words.shape[0]
    cap = max_parts if max_parts is not None else n + 1
    perm = np.zeros(max(n, 1), np.uint32)
    off = np.zeros(cap + 2, np.uint32)
    masks = np.zeros((cap + 1, 6), np.uint32)
    npart = C.c_uint32(0)
    check(load().tm_partition(_ptr(words), n, maxp, _ptr(perm), _ptr(off), _ptr(masks), cap, C.byref(npart)),
          "tm_partition")
    k = npart.value
    return perm[:n].copy(), off[: k + 1].copy(), masks[:k].copy()


class Fib:
    """Device FIB (tm_fib)."""

    def __init__(self, words, part_offsets, part_masks, key_off, keys, filter_ids=None, reserve=None):
        self.h = C.c_void_p()
        check(load().tm_fib_create(C.byref(self.h)), "tm_fib_create")
        if reserve is not None:
            self.set_reserve(reserve)
        self.upload(words, part_offsets, part_masks, key_off, keys, filter_ids)

    def upload(self, words, part_offsets, part_masks, key_off, keys, filter_ids=None):
        words = np.ascontiguousarray(words, np.uint32)
        po = np.ascontiguousarray(part_offsets, np.uint32)
        pm = np.ascontiguousarray(part_masks, np.uint32)
        ko = np.ascontiguousarray(key_off, np.uint32)
        kv = np.ascontiguousarray(keys, np.uint32)
        fid = None if filter_ids is None else np.ascontiguousarray(filter_ids, np.uint32)
        n = words.shape[0]
        check(load().tm_fib_upload(self.h, _ptr(words), n, _ptr(po), _ptr(pm), len(po) - 1, _ptr(ko), _ptr(kv),
                                   _ptr(fid)), "tm_fib_upload")

    @classmethod
    def empty(cls):
        """A FIB with no table yet (upload() or load() gives it one)."""
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        check(load().tm_fib_create(C.byref(self.h)), "tm_fib_create")
        return self

    @classmethod
    def from_image(cls, path):
        """A FIB holding the table of an image written by save()."""
        self = cls.empty()
        self.load(path)
        return self

    def set_reserve(self, percent):
        """Spare blocks per partition (percent) for the next upload."""
        check(load().tm_fib_set_reserve(self.h, percent), "tm_fib_set_reserve")

    def insert(self, words, key_off, keys, filter_ids=None, maxp=256 * 1024):
        """Insert filters into the resident table (tm_fib_insert)."""
        words = np.ascontiguousarray(words, np.uint32).reshape(-1, 6)
        ko = np.ascontiguousarray(key_off, np.uint32)
        kv = np.ascontiguousarray(keys, np.uint32)
        fid = None if filter_ids is None else np.ascontiguousarray(filter_ids, np.uint32)
        check(load().tm_fib_insert(self.h, _ptr(words), words.shape[0], _ptr(ko), _ptr(kv), _ptr(fid), maxp),
              "tm_fib_insert")

    def insert_live(self, words, key_off, keys, filter_ids=None):
        """Insert while submissions are in flight (tm_fib_insert_live);
        returns False (nothing inserted) when the batch needs tm_fib_insert."""
        words = np.ascontiguousarray(words, np.uint32).reshape(-1, 6)
        ko = np.ascontiguousarray(key_off, np.uint32)
        kv = np.ascontiguousarray(keys, np.uint32)
        fid = None if filter_ids is None else np.ascontiguousarray(filter_ids, np.uint32)
        rc = load().tm_fib_insert_live(self.h, _ptr(words), words.shape[0], _ptr(ko), _ptr(kv), _ptr(fid))
        if rc == TM_ERR_BUSY:
            return False
        check(rc, "tm_fib_insert_live")
        return True

    def save(self, path):
        """Write a versioned image of the resident table (tm_fib_save)."""
        check(load().tm_fib_save(self.h, os.fsencode(path)), "tm_fib_save")

    def load(self, path):
        """Replace the table with an image from save() (tm_fib_load)."""
        check(load().tm_fib_load(self.h, os.fsencode(path)), "tm_fib_load")

    def dictionary(self):
        """The sorted distinct keys that KEY_BITMAP rows index."""
        n = C.c_uint64()
        check(load().tm_fib_key_dictionary(self.h, None, 0, C.byref(n)), "tm_fib_key_dictionary")
        d = np.zeros(n.value, np.uint32)
        check(load().tm_fib_key_dictionary(self.h, _ptr(d), d.size, C.byref(n)), "tm_fib_key_dictionary")
        return d

    def stats(self):
        s = FibStats()
        check(load().tm_fib_stats_get(self.h, C.byref(s)), "tm_fib_stats_get")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.h:
            load().tm_fib_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_fib(words, key_off, keys, maxp, reserve=None, times=None):
    """Partition a canonical table on the GPU and upload it.

    Returns (fib, perm, part_offsets, masks); filter ids reported in HITS
    mode are positions in the canonical table.  `times`, a dict, receives
    the seconds of the partitioning ("partition") and of the rest
    ("upload")."""
    t0 = time.perf_counter()
    perm, off, masks = partition(words, maxp)
    t1 = time.perf_counter()
    pw = words[perm]
    cnt = (key_off[1:] - key_off[:-1]).astype(np.int64)[perm]
    ko = np.zeros(len(perm) + 1, np.uint64)
    np.cumsum(cnt, out=ko[1:])
    starts = key_off[:-1][perm].astype(np.int64)
    idx = np.repeat(starts - ko[:-1].astype(np.int64), cnt) + np.arange(int(ko[-1]), dtype=np.int64)
    pk = keys[idx] if len(idx) else np.zeros(0, np.uint32)
    fib = Fib(pw, off, masks, ko.astype(np.uint32), pk, perm.astype(np.uint32), reserve=reserve)
    if times is not None:
        times["partition"] = t1 - t0
        times["upload"] = time.perf_counter() - t1
    return fib, perm, off, masks


def zipf_runs(n, smin, smax, s=1.2, seed=1):
    """Offsets cutting n consecutive filters into runs of Zipf(s) sizes in
    [smin, smax] (a power law p(size) ~ size^-s).  The last run takes the
    remainder; a remainder below smin joins the run before it, and when
    that would pass smax the two are split in half (both then lie in
    [smin, smax] as long as smax >= 2 * smin)."""
    rng = np.random.default_rng(seed)
    off = [0]
    a = 1.0 - s
    lo, hi = smin ** a, smax ** a
    while off[-1] < n:
        u = rng.random()
        size = int(round((lo + u * (hi - lo)) ** (1.0 / a)))
        off.append(min(n, off[-1] + max(size, 1)))
    if len(off) > 2 and off[-1] - off[-2] < smin:
        off.pop(-2)
        if off[-1] - off[-2] > smax:
            off.insert(-1, off[-2] + (off[-1] - off[-2]) // 2)
    return np.array(off, np.uint32)


def fib_from_runs(words, key_off, keys, offsets, reserve=None):
    """Upload a table whose partitions are the given runs of consecutive
    filters; each mask is the AND of its run (all-zero: selected by every
    query).  Filter ids are positions in `words`."""
    words = np.ascontiguousarray(words, np.uint32)
    off = np.asarray(offsets, np.int64)
    masks = np.stack([np.bitwise_and.reduce(words[off[i]:off[i + 1]], axis=0) if off[i + 1] > off[i]
                      else np.full(6, 0xFFFFFFFF, np.uint32) for i in range(len(off) - 1)]).astype(np.uint32)
    fid = np.arange(words.shape[0], dtype=np.uint32)
    return Fib(words, off.astype(np.uint32), masks, np.asarray(key_off).astype(np.uint32), keys, fid,
               reserve=reserve), masks


def uneven_units(n_parts, bmin=1, bmax=4096, seed=2):
    """Batch sizes per partition, log-uniform in [bmin, bmax]; returns
    (unit_part, unit_off) for Ring.submit_units with one unit per partition."""
    rng = np.random.default_rng(seed)
    b = np.exp(rng.uniform(np.log(bmin), np.log(bmax + 1), n_parts)).astype(np.int64)
    b = np.clip(b, bmin, bmax)
    uo = np.zeros(n_parts + 1, np.uint64)
    np.cumsum(b, out=uo[1:])
    return np.arange(n_parts, dtype=np.uint32), uo.astype(np.uint32)


def config(mode=MODE_KEYS_UNIQUE, schedule=SCHED_PERSISTENT, **kw):
    c = MatchConfig()
    load().tm_match_config_default(C.byref(c))
    c.mode = mode
    c.schedule = schedule
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Ring:
    """One in-flight submission context (tm_ring)."""

    def __init__(self, fib, max_queries, cfg=None, **kw):
        self.fib = fib
        self.cfg = cfg if cfg is not None else config(**kw)
        self.h = C.c_void_p()
        self.max_queries = max_queries
        check(load().tm_ring_create(fib.h, max_queries, C.byref(self.cfg), C.byref(self.h)), "tm_ring_create")
        self.nq = 0

    def configure(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else config(**kw)
        check(load().tm_ring_config(self.h, C.byref(self.cfg)), "tm_ring_config")

    def submit(self, queries):
        q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 6)
        self.nq = q.shape[0]
        check(load().tm_submit(self.h, _ptr(q), self.nq), "tm_submit")

    def submit_units(self, queries, unit_part, unit_off):
        """Explicit (partition, query range) units, no lookup (tm_submit_units);
        queries None: the resident ones."""
        up = np.ascontiguousarray(unit_part, np.uint32)
        uo = np.ascontiguousarray(unit_off, np.uint32)
        nq = int(uo[-1]) if len(uo) else 0
        q = None if queries is None else np.ascontiguousarray(queries, np.uint32).reshape(-1, 6)
        self.nq = nq
        check(load().tm_submit_units(self.h, _ptr(q), nq, _ptr(up), _ptr(uo), len(up)), "tm_submit_units")

    def submit_resident(self, nq):
        self.nq = nq
        check(load().tm_submit_resident(self.h, nq), "tm_submit_resident")

    def poll(self):
        rc = load().tm_poll(self.h)
        if rc == TM_ERR_BUSY:
            return False
        check(rc, "tm_poll")
        return True

    def wait(self):
        st = Status()
        check(load().tm_wait(self.h, C.byref(st)), "tm_wait")
        return st

    def results(self):
        off = np.zeros(self.nq + 1, np.uint64)
        check(load().tm_read_results(self.h, _ptr(off), C.c_void_p(0), 0), "tm_read_results")
        vals = np.zeros(int(off[-1]), np.uint32)
        check(load().tm_read_results(self.h, _ptr(off), _ptr(vals), vals.size), "tm_read_results")
        return off, vals

    def key_bitmaps(self):
        """Key bitmap rows of the finished submission: (nq, kw) uint32."""
        kw = C.c_uint32()
        rc = load().tm_read_key_bitmaps(self.h, None, 0, C.byref(kw))
        if rc not in (TM_OK, TM_ERR_OVERFLOW):
            check(rc, "tm_read_key_bitmaps")
        rows = np.zeros((self.nq, kw.value), np.uint32)
        check(load().tm_read_key_bitmaps(self.h, _ptr(rows), rows.size, C.byref(kw)), "tm_read_key_bitmaps")
        return rows

    def set_host_rows(self, buf):
        """Deliver each KEY_BITMAP submission's rows into `buf` (a PinnedRows
        or None to stop)."""
        self.host_rows = buf
        if buf is None:
            check(load().tm_ring_set_host_rows(self.h, None, 0), "tm_ring_set_host_rows")
        else:
            check(load().tm_ring_set_host_rows(self.h, C.c_void_p(buf.ptr), buf.words), "tm_ring_set_host_rows")

    def set_host_rows_packed(self, buf):
        """Deliver each KEY_BITMAP submission packed into `buf` (a
        PackedRows or None to stop): per query TM_ROW_FULL or its row's
        index, and the rows that are not the whole dictionary."""
        self.host_rows = buf
        if buf is None:
            check(load().tm_ring_set_host_rows_packed(self.h, None, 0, None, 0), "tm_ring_set_host_rows_packed")
        else:
            check(load().tm_ring_set_host_rows_packed(self.h, C.c_void_p(buf.row_of.ptr), buf.row_of.words,
                                                      C.c_void_p(buf.rows.ptr), buf.rows.words),
                  "tm_ring_set_host_rows_packed")

    def lookup_counts(self):
        n = int(self.fib.stats()["n_parts"])
        c = np.zeros(max(n, 1), np.uint32)
        check(load().tm_ring_lookup_counts(self.h, _ptr(c), n), "tm_ring_lookup_counts")
        return c[:n]

    def autotune(self, nq, candidates, reps=1):
        """Pick the fastest of `candidates` (MatchConfig or dicts of config
        fields) on the nq resident queries; returns (index, times_ms)."""
        cands = [c if isinstance(c, MatchConfig) else config(mode=self.cfg.mode, **c) for c in candidates]
        arr = (MatchConfig * len(cands))(*cands)
        best = C.c_uint32()
        times = np.zeros(len(cands), np.float64)
        check(load().tm_ring_autotune(self.h, nq, arr, len(cands), reps, C.byref(best), _ptr(times)),
              "tm_ring_autotune")
        self.cfg = cands[best.value]
        return best.value, times

    def set_bucket_segments(self, seg):
        a = np.ascontiguousarray(seg, np.uint32)
        check(load().tm_ring_set_bucket_segments(self.h, _ptr(a), a.size), "tm_ring_set_bucket_segments")

    def bucket_segments(self):
        a = np.zeros(SIZE_BUCKETS, np.uint32)
        check(load().tm_ring_get_bucket_segments(self.h, _ptr(a), a.size), "tm_ring_get_bucket_segments")
        return a

    def autotune_buckets(self, unit_part, unit_off, candidates=(4, 16, 64, 256, 1024, 1 << 20), reps=1):
        """Tune blocks per work item by partition-size bucket on the resident
        queries' units (tm_ring_autotune_buckets); returns (base_ms, best_ms,
        table)."""
        up = np.ascontiguousarray(unit_part, np.uint32)
        uo = np.ascontiguousarray(unit_off, np.uint32)
        c = np.ascontiguousarray(candidates, np.uint32)
        b0, b1 = C.c_double(), C.c_double()
        check(load().tm_ring_autotune_buckets(self.h, _ptr(up), _ptr(uo), len(up), _ptr(c), c.size, reps,
                                              C.byref(b0), C.byref(b1)), "tm_ring_autotune_buckets")
        self.nq = int(uo[-1])
        return b0.value, b1.value, self.bucket_segments()

    def unit_log(self):
        """(features [n,16] f32, hits [n], keys [n], ticks [n], units [n,3]
        = partition, batch, queries) of the last plain submission made with
        log_units=1."""
        n = C.c_uint32()
        check(load().tm_ring_unit_log(self.h, None, None, None, None, None, 0, C.byref(n)), "tm_ring_unit_log")
        k = n.value
        f = np.zeros((k, UNIT_FEATURES), np.float32)
        h = np.zeros(k, np.uint32)
        kk = np.zeros(k, np.uint32)
        t = np.zeros(k, np.uint32)
        u = np.zeros((k, 3), np.uint32)
        check(load().tm_ring_unit_log(self.h, _ptr(f), _ptr(h), _ptr(kk), _ptr(t), _ptr(u), k, C.byref(n)),
              "tm_ring_unit_log")
        return f, h, kk, t, u

    def set_cost_model(self, mlp, flags=None):
        """Price PERSISTENT submissions with `mlp` (None: stop)."""
        if flags is None:
            flags = COST_PRESIZE | COST_ORDER
        self.mlp = mlp
        check(load().tm_ring_set_cost_model(self.h, mlp.h if mlp is not None else None,
                                            flags if mlp is not None else 0), "tm_ring_set_cost_model")

    def query_buffer(self):
        p = C.c_void_p()
        check(load().tm_ring_query_buffer(self.h, C.byref(p)), "tm_ring_query_buffer")
        return p.value

    def close(self):
        if self.h:
            load().tm_ring_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory from tm_dev_alloc, with synchronous copies."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(load().tm_dev_alloc(C.byref(p), int(nbytes)), "tm_dev_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(load().tm_dev_upload(C.c_void_p(self.ptr), _ptr(a), a.nbytes), "tm_dev_upload")

    def download(self, a):
        assert a.nbytes <= self.nbytes and a.flags["C_CONTIGUOUS"]
        check(load().tm_dev_download(_ptr(a), C.c_void_p(self.ptr), a.nbytes), "tm_dev_download")
        return a

    def close(self):
        if getattr(self, "ptr", None):
            load().tm_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedRows:
    """Pinned host memory (tm_host_alloc) viewed as uint32 words."""

    def __init__(self, words):
        p = C.c_void_p()
        check(load().tm_host_alloc(C.byref(p), int(words) * 4), "tm_host_alloc")
        self.ptr = p.value
        self.words = int(words)
        self.array = np.ctypeslib.as_array(C.cast(p, U32P), shape=(max(self.words, 1),))[:self.words]

    def rows(self, nq, kw):
        return self.array[:nq * kw].reshape(nq, kw)

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            load().tm_host_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackedRows:
    """Pinned host buffers of packed KEY_BITMAP delivery for up to nq
    queries of kw-word rows."""

    def __init__(self, nq, kw):
        self.kw = int(kw)
        self.row_of = PinnedRows(nq)
        self.rows = PinnedRows(int(nq) * self.kw)

    def unpack(self, nq, n_rows, n_keys):
        """The full (nq, kw) rows of a delivered submission."""
        row_of = self.row_of.array[:nq]
        packed = self.rows.array[:n_rows * self.kw].reshape(n_rows, self.kw)
        full = np.zeros(self.kw, np.uint32)
        full[:n_keys // 32] = 0xFFFFFFFF
        if n_keys % 32:
            full[n_keys // 32] = (1 << (n_keys % 32)) - 1
        out = np.empty((nq, self.kw), np.uint32)
        is_full = row_of == ROW_FULL
        out[is_full] = full
        out[~is_full] = packed[row_of[~is_full]]
        return out

    def close(self):
        self.row_of.close()
        self.rows.close()


def expand_key_rows_packed(row_of, rows, dictionary, threads=0):
    """Packed delivery (row_of [nq], rows [n_rows, kw]) -> KEYS_UNIQUE
    (offsets, values), expanded on the host (tm_expand_key_rows_packed)."""
    row_of = np.ascontiguousarray(row_of, np.uint32)
    rows = np.ascontiguousarray(rows, np.uint32)
    n_rows, kw = rows.shape
    nq = row_of.size
    d = np.ascontiguousarray(dictionary, np.uint32)
    off = np.zeros(nq + 1, np.uint64)
    lib = load()
    check(lib.tm_expand_key_rows_packed(_ptr(row_of), _ptr(rows), n_rows, nq, kw, _ptr(d), d.size, _ptr(off), None,
                                        0, threads), "tm_expand_key_rows_packed")
    vals = np.zeros(int(off[-1]), np.uint32)
    check(lib.tm_expand_key_rows_packed(_ptr(row_of), _ptr(rows), n_rows, nq, kw, _ptr(d), d.size, _ptr(off),
                                        _ptr(vals), vals.size, threads), "tm_expand_key_rows_packed")
    return off, vals


def expand_key_bitmaps(rows, dictionary, threads=0):
    """KEY_BITMAP rows (nq, kw) over `dictionary` -> KEYS_UNIQUE (offsets,
    values), expanded on the host (tm_expand_key_bitmaps)."""
    rows = np.ascontiguousarray(rows, np.uint32)
    nq, kw = rows.shape
    d = np.ascontiguousarray(dictionary, np.uint32)
    off = np.zeros(nq + 1, np.uint64)
    lib = load()
    check(lib.tm_expand_key_bitmaps(_ptr(rows), nq, kw, _ptr(d), d.size, _ptr(off), None, 0, threads),
          "tm_expand_key_bitmaps")
    vals = np.zeros(int(off[-1]), np.uint32)
    check(lib.tm_expand_key_bitmaps(_ptr(rows), nq, kw, _ptr(d), d.size, _ptr(off), _ptr(vals), vals.size, threads),
          "tm_expand_key_bitmaps")
    return off, vals


def match(fib, queries, mode=MODE_KEYS_UNIQUE, schedule=SCHED_PERSISTENT, **kw):
    """Synchronous match: returns (offsets, values, status dict)."""
    q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 6)
    r = Ring(fib, max(q.shape[0], 1), mode=mode, schedule=schedule, **kw)
    try:
        r.submit(q)
        st = r.wait()
        off, vals = r.results()
    finally:
        r.close()
    return off, vals, st.as_dict()


# ------------------------------------------------------------------ serving

class Server:
    """Continuous-batching serving loop (tm_server)."""

    def __init__(self, fib, n_slots=4, slot_queries=16384, hbm_budget=0, cfg=None, flush_queries=0, flush_us=0,
                 step_graphs=0, **kw):
        sc = ServerConfig()
        load().tm_server_config_default(C.byref(sc))
        sc.n_slots = n_slots
        sc.slot_queries = slot_queries
        sc.hbm_budget = hbm_budget
        sc.flush_queries = flush_queries
        sc.flush_us = flush_us
        sc.step_graphs = step_graphs
        self.cfg = cfg if cfg is not None else config(**kw)
        self.fib = fib
        self.h = C.c_void_p()
        check(load().tm_server_create(fib.h, C.byref(sc), C.byref(self.cfg), C.byref(self.h)), "tm_server_create")

    def submit(self, queries):
        q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 6)
        rid = C.c_uint64()
        check(load().tm_server_submit(self.h, _ptr(q), q.shape[0], C.byref(rid)), "tm_server_submit")
        return rid.value

    def insert(self, words, key_off, keys, filter_ids=None, maxp=256 * 1024):
        w = np.ascontiguousarray(words, np.uint32).reshape(-1, 6)
        ko = np.ascontiguousarray(key_off, np.uint32)
        kv = np.ascontiguousarray(keys, np.uint32)
        fid = None if filter_ids is None else np.ascontiguousarray(filter_ids, np.uint32)
        rid = C.c_uint64()
        check(load().tm_server_insert(self.h, _ptr(w), w.shape[0], _ptr(ko), _ptr(kv), _ptr(fid), maxp,
                                      C.byref(rid)), "tm_server_insert")
        return rid.value

    def wait(self, rid, timeout_ms=-1.0):
        rc = load().tm_server_wait(self.h, rid, timeout_ms)
        if rc == TM_ERR_BUSY:
            return False
        check(rc, "tm_server_wait")
        return True

    def collect(self, cap=65536):
        ids = np.zeros(cap, np.uint64)
        n = C.c_uint32()
        check(load().tm_server_collect(self.h, _ptr(ids), cap, C.byref(n)), "tm_server_collect")
        return ids[:n.value]

    def info(self, rid):
        inf = RequestInfo()
        check(load().tm_server_result(self.h, rid, C.byref(inf), None, None, 0), "tm_server_result")
        return inf

    def result(self, rid, release=True):
        """(info, offsets, values) of a completed request."""
        inf = self.info(rid)
        off = np.zeros(inf.n_queries + 1, np.uint64)
        vals = np.zeros(inf.n_values, np.uint32)
        check(load().tm_server_result(self.h, rid, None, _ptr(off), _ptr(vals), vals.size), "tm_server_result")
        if release:
            self.release(rid)
        return inf, off, vals

    def release(self, rid):
        check(load().tm_server_release(self.h, rid), "tm_server_release")

    def set_cost_model(self, mlp, flags=None):
        if flags is None:
            flags = COST_PRESIZE | COST_ORDER
        check(load().tm_server_set_cost_model(self.h, mlp.h if mlp is not None else None,
                                              flags if mlp is not None else 0), "tm_server_set_cost_model")

    def stats(self):
        st = ServerStats()
        check(load().tm_server_stats_get(self.h, C.byref(st)), "tm_server_stats_get")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            load().tm_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ cost model

UNIT_FEATURES = 16
COST_PRESIZE = 1
COST_ORDER = 2


def encode_targets(hits, keys, ticks):
    """Training targets of the cost model: log2(1 + x) / 16 of (hits, key
    instances, ticks)."""
    return np.stack([np.log2(1.0 + np.asarray(x, np.float64)) / 16.0 for x in (hits, keys, ticks)],
                    axis=1).astype(np.float32)


def decode_targets(y):
    """Inverse of encode_targets: (hits, keys, ticks) as float64."""
    y = np.maximum(np.asarray(y, np.float64), 0.0)
    return tuple(np.exp2(16.0 * y[:, i]) - 1.0 for i in range(3))


class MLP:
    """n_in -> h -> h -> n_out FP32 perceptron on the GPU (tm_mlp)."""

    def __init__(self, n_in=UNIT_FEATURES, n_hidden=64, n_out=3, seed=1):
        self.shape = (n_in, n_hidden, n_out)
        self.h = C.c_void_p()
        check(load().tm_mlp_create(n_in, n_hidden, n_out, seed, C.byref(self.h)), "tm_mlp_create")
        self.n_params = int(load().tm_mlp_param_count(self.h))

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        check(load().tm_mlp_get_params(self.h, _ptr(p), self.n_params), "tm_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(load().tm_mlp_set_params(self.h, _ptr(p), p.size), "tm_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.shape[0])
        y = np.zeros((x.shape[0], self.shape[2]), np.float32)
        check(load().tm_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(y)), "tm_mlp_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.shape[0])
        t = np.ascontiguousarray(t, np.float32).reshape(-1, self.shape[2])
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        check(load().tm_mlp_loss_grad(self.h, _ptr(x), _ptr(t), x.shape[0], C.byref(loss), _ptr(g)),
              "tm_mlp_loss_grad")
        return float(loss.value), g

    def train(self, x, t, epochs=10, batch=1024, lr=1e-3, seed=7):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.shape[0])
        t = np.ascontiguousarray(t, np.float32).reshape(-1, self.shape[2])
        loss = C.c_float()
        check(load().tm_mlp_train(self.h, _ptr(x), _ptr(t), x.shape[0], epochs, batch, lr, seed,
                                  C.byref(loss)), "tm_mlp_train")
        return float(loss.value)

    def calibrate(self, x, *targets):
        """Shift each output's bias so that the decoded predictions sum to
        the measured totals on (x, targets): a model fitted in log space
        predicts geometric means, which under-estimate sums of skewed
        quantities (hits, keys, ticks) that buffer sizing needs."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.shape[0])
        y = self.forward(x).astype(np.float64)
        p = self.params()
        n_in, h, n_out = self.shape
        b3 = n_in * h + h + h * h + h + h * n_out
        for i, t in enumerate(targets[:n_out]):
            pred = np.exp2(16.0 * np.maximum(y[:, i], 0.0)) - 1.0
            tot, ptot = float(np.sum(t)), float(pred.sum())
            if tot > 0 and ptot > 0:
                p[b3 + i] += np.float32(np.log2(tot / ptot) / 16.0)
        self.set_params(p)

    def close(self):
        if self.h:
            load().tm_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mape(pred, true, floor=1.0):
    """Mean absolute percentage error over entries with true >= floor."""
    pred = np.asarray(pred, np.float64)
    true = np.asarray(true, np.float64)
    m = true >= floor
    return float(np.mean(np.abs(pred[m] - true[m]) / true[m])) if m.any() else 0.0


def split(off, vals):
    """Per-query lists from (offsets, values)."""
    return [vals[int(off[i]):int(off[i + 1])] for i in range(len(off) - 1)]


# ------------------------------------------------------------------ file I/O

def read_fib_file(path, fmt=IO_BINARY):
    """Filter table in the reference's formats (tm_io_read_fib):
    (words n x 6, key_offsets n + 1, keys, partition ids)."""
    nf, nk = C.c_uint64(), C.c_uint64()
    bp = os.fsencode(path)
    check(load().tm_io_scan_fib(bp, fmt, C.byref(nf), C.byref(nk)), "tm_io_scan_fib")
    words = np.zeros((nf.value, 6), np.uint32)
    ko = np.zeros(nf.value + 1, np.uint32)
    keys = np.zeros(nk.value, np.uint32)
    parts = np.zeros(nf.value, np.uint32)
    check(load().tm_io_read_fib(bp, fmt, _ptr(words), _ptr(ko), _ptr(keys), _ptr(parts), nf.value, nk.value),
          "tm_io_read_fib")
    return words, ko, keys, parts


def write_fib_file(path, words, key_offsets, keys, fmt=IO_BINARY, partitions=None):
    w = np.ascontiguousarray(words, np.uint32).reshape(-1, 6)
    ko = np.ascontiguousarray(key_offsets, np.uint32)
    kv = np.ascontiguousarray(keys, np.uint32)
    pt = None if partitions is None else np.ascontiguousarray(partitions, np.uint32)
    check(load().tm_io_write_fib(os.fsencode(path), fmt, _ptr(w), w.shape[0], _ptr(ko), _ptr(kv), _ptr(pt)),
          "tm_io_write_fib")


def read_query_file(path, fmt=IO_BINARY):
    """Queries in the reference's formats (tm_io_read_queries): n x 6 words."""
    n = C.c_uint64()
    bp = os.fsencode(path)
    check(load().tm_io_scan_queries(bp, fmt, C.byref(n)), "tm_io_scan_queries")
    q = np.zeros((n.value, 6), np.uint32)
    check(load().tm_io_read_queries(bp, fmt, _ptr(q), n.value), "tm_io_read_queries")
    return q


def write_query_file(path, words, fmt=IO_BINARY):
    q = np.ascontiguousarray(words, np.uint32).reshape(-1, 6)
    check(load().tm_io_write_queries(os.fsencode(path), fmt, _ptr(q), q.shape[0]), "tm_io_write_queries")


def read_prefix_file(path, fmt=IO_BINARY):
    """Partition prefixes (tm_io_read_prefixes): dict of words (n x 6),
    length, partition, size."""
    n = C.c_uint64()
    bp = os.fsencode(path)
    check(load().tm_io_scan_prefixes(bp, fmt, C.byref(n)), "tm_io_scan_prefixes")
    arr = (Prefix * max(n.value, 1))()
    check(load().tm_io_read_prefixes(bp, fmt, arr, n.value), "tm_io_read_prefixes")
    a = np.ctypeslib.as_array(C.cast(arr, U32P), shape=(max(n.value, 1) * 9,)).reshape(-1, 9)[:n.value]
    return {"words": a[:, :6].copy(), "length": a[:, 6].copy(), "partition": a[:, 7].copy(), "size": a[:, 8].copy()}


def write_prefix_file(path, words, length, partition, size, fmt=IO_BINARY):
    a = np.zeros((len(length), 9), np.uint32)
    a[:, :6] = np.asarray(words, np.uint32).reshape(-1, 6)
    a[:, 6] = length
    a[:, 7] = partition
    a[:, 8] = size
    check(load().tm_io_write_prefixes(os.fsencode(path), fmt, _ptr(a), a.shape[0]), "tm_io_write_prefixes")


def convert_file(src, src_fmt, dst, dst_fmt, kind=KIND_FIB):
    """The reference's convert_workload (tm_io_convert); returns records written."""
    n = C.c_uint64()
    check(load().tm_io_convert(os.fsencode(src), src_fmt, os.fsencode(dst), dst_fmt, kind, C.byref(n)),
          "tm_io_convert")
    return n.value


def sample_file(src, fmt, dst, kind, n_total, k, seed):
    """The reference's sample_workload (tm_io_sample); returns records written."""
    n = C.c_uint64()
    check(load().tm_io_sample(os.fsencode(src), fmt, os.fsencode(dst), kind, n_total, k, seed, C.byref(n)),
          "tm_io_sample")
    return n.value


def split_file(src, fmt, prefixes_path, filters_path, max_size=100, min_k=2):
    """The reference's split_workload (tm_io_split); returns partitions."""
    n = C.c_uint64()
    pp = os.fsencode(prefixes_path) if prefixes_path else None
    fp = os.fsencode(filters_path) if filters_path else None
    check(load().tm_io_split(os.fsencode(src), fmt, pp, fp, max_size, min_k, C.byref(n)), "tm_io_split")
    return n.value
