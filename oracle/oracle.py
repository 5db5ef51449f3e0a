"""ctypes wrapper of liboracle.so, the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The engine never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

P = C.c_void_p
_lib = None


def build(reference=True):
    """Compile liboracle.so, and the reference programs into _ref/ when the
    reference sources are present."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    if reference and os.path.isdir(os.environ.get("REF", "/root/reference")):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run oracle.build()")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_subset": (C.c_int, [P, P]),
        "orc_compare": (C.c_int, [P, P]),
        "orc_leftmost_diff": (C.c_int, [P, P]),
        "orc_from_string": (None, [C.c_char_p, P]),
        "orc_consolidate": (C.c_int, [P, C.c_uint64, C.POINTER(P), C.POINTER(C.c_uint64), C.POINTER(P), C.POINTER(P)]),
        "orc_match": (C.c_int, [P, C.c_uint64, P, P, P, C.c_uint64, C.c_int, C.c_int, P, C.POINTER(P)]),
        "orc_partition": (C.c_int, [P, C.c_uint32, C.c_uint32, P, P, P, C.c_uint32, C.POINTER(C.c_uint32)]),
        "orc_lookup_build": (P, [P, C.c_uint32]),
        "orc_lookup_query": (C.c_uint32, [P, P, P]),
        "orc_lookup_free": (None, [P]),
        "orc_schedule_count": (C.c_uint64, [P, P, P, C.c_uint32, P, C.c_uint64, C.c_int]),
        "orc_bruteforce_count": (C.c_uint64, [P, C.c_uint64, P, C.c_uint64, C.c_int]),
        "orc_free": (None, [P]),
        "orc_schedule_match": (C.c_int, [P, P, P, P, P, C.c_uint32, P, C.c_uint64, C.c_int, C.c_int, P,
                                         C.POINTER(P)]),
        "orc_compiler": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


def _w(x):
    return np.ascontiguousarray(x, np.uint32)


def from_string(s):
    out = np.zeros(6, np.uint32)
    load().orc_from_string(s.encode(), _p(out))
    return out


def subset(f, q):
    return bool(load().orc_subset(_p(_w(f)), _p(_w(q))))


def compare(a, b):
    return load().orc_compare(_p(_w(a)), _p(_w(b)))


def leftmost_diff(a, b):
    return load().orc_leftmost_diff(_p(_w(a)), _p(_w(b)))


def consolidate(ops):
    """ops: (n, 8) uint32 rows = 6 words, key, op (0 add, 1 remove)."""
    ops = _w(ops).reshape(-1, 8)
    w, ko, k = C.c_void_p(), C.c_void_p(), C.c_void_p()
    n = C.c_uint64()
    lib = load()
    lib.orc_consolidate(_p(ops), ops.shape[0], C.byref(w), C.byref(n), C.byref(ko), C.byref(k))
    m = n.value
    words = np.ctypeslib.as_array(C.cast(w, C.POINTER(C.c_uint32)), shape=(max(m * 6, 1),))[: m * 6].copy().reshape(m, 6)
    off = np.ctypeslib.as_array(C.cast(ko, C.POINTER(C.c_uint64)), shape=(m + 1,)).copy()
    nk = int(off[-1])
    keys = np.ctypeslib.as_array(C.cast(k, C.POINTER(C.c_uint32)), shape=(max(nk, 1),))[:nk].copy()
    for x in (w, ko, k):
        lib.orc_free(x)
    return words, off, keys


def match(words, key_off, keys, queries, mode=0, threads=None):
    """Brute force; mode 0 unique keys, 1 all keys (sorted), 2 filter ids."""
    words = _w(words).reshape(-1, 6)
    queries = _w(queries).reshape(-1, 6)
    ko = np.ascontiguousarray(key_off, np.uint64)
    kv = _w(keys)
    nq = queries.shape[0]
    off = np.zeros(nq + 1, np.uint64)
    vals = C.c_void_p()
    t = threads or min(16, os.cpu_count() or 1)
    load().orc_match(_p(words), words.shape[0], _p(ko), _p(kv), _p(queries), nq, mode, t, _p(off), C.byref(vals))
    n = int(off[-1])
    v = np.ctypeslib.as_array(C.cast(vals, C.POINTER(C.c_uint32)), shape=(max(n, 1),))[:n].copy()
    load().orc_free(vals)
    return off, v


def schedule_match(words, key_off, keys, part_off, masks, queries, mode=0, threads=None):
    """The reference schedule (lookup + per-partition brute force) with key
    output; words/key_off/keys in partitioned order."""
    words = _w(words).reshape(-1, 6)
    queries = _w(queries).reshape(-1, 6)
    m = _w(masks).reshape(-1, 6)
    po = _w(part_off)
    ko = np.ascontiguousarray(key_off, np.uint64)
    nq = queries.shape[0]
    off = np.zeros(nq + 1, np.uint64)
    vals = C.c_void_p()
    t = threads or min(16, os.cpu_count() or 1)
    load().orc_schedule_match(_p(words), _p(ko), _p(_w(keys)), _p(po), _p(m), m.shape[0], _p(queries), nq, mode, t,
                              _p(off), C.byref(vals))
    n = int(off[-1])
    v = np.ctypeslib.as_array(C.cast(vals, C.POINTER(C.c_uint32)), shape=(max(n, 1),))[:n].copy()
    load().orc_free(vals)
    return off, v


def compiler():
    return load().orc_compiler().decode()


def partition(words, maxp):
    """Returns (rc, perm, part_offsets, masks)."""
    words = _w(words).reshape(-1, 6)
    n = words.shape[0]
    perm = np.zeros(max(n, 1), np.uint32)
    off = np.zeros(n + 2, np.uint32)
    masks = np.zeros((n + 1, 6), np.uint32)
    npart = C.c_uint32()
    rc = load().orc_partition(_p(words), n, maxp, _p(perm), _p(off), _p(masks), n + 1, C.byref(npart))
    k = npart.value
    return rc, perm[:n].copy(), off[: k + 1].copy(), masks[:k].copy()


class Lookup:
    def __init__(self, masks):
        self.masks = _w(masks).reshape(-1, 6)
        self.h = load().orc_lookup_build(_p(self.masks), self.masks.shape[0])

    def query(self, q):
        sel = np.zeros(max(self.masks.shape[0], 1), np.uint32)
        n = load().orc_lookup_query(self.h, _p(_w(q)), _p(sel))
        return sel[:n].copy()

    def __del__(self):
        if getattr(self, "h", None):
            load().orc_lookup_free(self.h)
            self.h = None


def schedule_count(words, part_off, masks, queries, threads=None):
    t = threads or min(16, os.cpu_count() or 1)
    q = _w(queries).reshape(-1, 6)
    m = _w(masks).reshape(-1, 6)
    return load().orc_schedule_count(_p(_w(words)), _p(_w(part_off)), _p(m), m.shape[0], _p(q), q.shape[0], t)


def bruteforce_count(words, queries, threads=None):
    t = threads or min(16, os.cpu_count() or 1)
    w = _w(words).reshape(-1, 6)
    q = _w(queries).reshape(-1, 6)
    return load().orc_bruteforce_count(_p(w), w.shape[0], _p(q), q.shape[0], t)


# ----------------------------------------------------------------- reference programs

def ref_available(name):
    return os.path.exists(os.path.join(REF_DIR, name))


def ref_test_matcher(words, key_off, keys, queries, workdir):
    """Run the reference test_matcher (oracle/_ref) on a table; returns the
    per-query key lists in the order the reference prints them."""
    fpath = os.path.join(workdir, "filters.bin")
    qpath = os.path.join(workdir, "queries.bin")
    with open(fpath, "wb") as f:
        for i in range(words.shape[0]):
            ks = keys[int(key_off[i]):int(key_off[i + 1])]
            f.write(np.uint32(0).tobytes())           # partition id (ignored)
            f.write(_w(words[i]).tobytes())
            f.write(np.uint32(len(ks)).tobytes())
            f.write(_w(ks).tobytes())
    _w(queries).tofile(qpath)
    out = subprocess.run([os.path.join(REF_DIR, "test_matcher"), f"F={fpath}", f"Q={qpath}"],
                         check=True, stdout=subprocess.PIPE, text=True).stdout
    res = []
    for line in out.splitlines():
        if not line.startswith("query="):
            continue
        parts = line.split()
        res.append(np.array([int(x) for x in parts[1:]], np.uint32))
    return res


def ref_partition(words, maxp, workdir, threads=1):
    """Run the reference partitioner; returns (order, part_of_position, masks
    by reference id, sizes by reference id)."""
    fpath = os.path.join(workdir, "pfilters.bin")
    opath = os.path.join(workdir, "pout.bin")
    _w(words).tofile(fpath)
    subprocess.run([os.path.join(REF_DIR, "ref_partition"), fpath, str(maxp), str(threads), opath], check=True)
    raw = np.fromfile(opath, np.uint8)
    n = int(raw[:4].view(np.uint32)[0])
    body = raw[4:4 + 8 * n].view(np.uint32).reshape(n, 2)
    rest = raw[4 + 8 * n:]
    np_ = int(rest[:4].view(np.uint32)[0])
    recs = rest[4:4 + np_ * 32].reshape(np_, 32)
    masks = recs[:, :24].copy().view(np.uint32).reshape(np_, 6)
    ids = recs[:, 24:28].copy().view(np.uint32).reshape(np_)
    sizes = recs[:, 28:32].copy().view(np.uint32).reshape(np_)
    m = np.zeros((np_, 6), np.uint32)
    s = np.zeros(np_, np.uint32)
    m[ids] = masks
    s[ids] = sizes
    return body[:, 0].copy(), body[:, 1].copy(), m, s


def ref_bitvector(pairs, workdir):
    """Run the reference bitvector<192> on (a, b) pairs of 6-word filters."""
    ipath = os.path.join(workdir, "pairs.bin")
    opath = os.path.join(workdir, "bv.bin")
    _w(pairs).reshape(-1, 12).tofile(ipath)
    subprocess.run([os.path.join(REF_DIR, "ref_bitvector"), ipath, opath], check=True)
    raw = np.fromfile(opath, np.uint8).reshape(-1, 36)
    return {
        "subset": raw[:, 0].astype(bool),
        "less": raw[:, 1].astype(bool),
        "greater": raw[:, 2].astype(bool),
        "leftmost_diff": raw[:, 3].astype(int),
        "popcount": raw[:, 4].astype(int),
        "next_bit": raw[:, 8:12].copy().view(np.uint32).ravel(),
        "words": raw[:, 12:36].copy().view(np.uint32).reshape(-1, 6),
    }


class RefPatricia:
    """The reference compact Patricia-trie matcher (_ref/libref_patricia.so,
    built from compact_patricia_predicate.hh and key_array.cc) over one
    table: match_unique results on the host's threads."""

    _lib = None

    def __init__(self, words, key_off, keys):
        if RefPatricia._lib is None:
            lib = C.CDLL(os.path.join(REF_DIR, "libref_patricia.so"))
            lib.rp_build.restype = P
            lib.rp_build.argtypes = [P, C.c_uint64, P, P]
            lib.rp_nodes.restype = C.c_uint64
            lib.rp_nodes.argtypes = [P]
            lib.rp_match_unique.restype = C.c_uint64
            lib.rp_match_unique.argtypes = [P, P, C.c_uint64, C.c_int, P, P, P]
            lib.rp_free.restype = None
            lib.rp_free.argtypes = [P]
            RefPatricia._lib = lib
        w = _w(words).reshape(-1, 6)
        ko = np.ascontiguousarray(key_off, np.uint64)
        kv = _w(keys)
        self.h = RefPatricia._lib.rp_build(_p(w), w.shape[0], _p(ko), _p(kv))
        if not self.h:
            raise ValueError("reference Patricia trie: empty table or a filter with more than 65535 keys")
        self.n_nodes = int(RefPatricia._lib.rp_nodes(self.h))

    def count(self, queries, threads=1):
        """Distinct keys per query (and their total), keys not kept."""
        q = _w(queries).reshape(-1, 6)
        cnt = np.zeros(q.shape[0], np.uint32)
        tot = RefPatricia._lib.rp_match_unique(self.h, _p(q), q.shape[0], threads, _p(cnt), None, None)
        return cnt, int(tot)

    def match_unique(self, queries, threads=1):
        """(offsets, keys): each query's distinct keys, sorted."""
        q = _w(queries).reshape(-1, 6)
        cnt, tot = self.count(q, threads)
        off = np.zeros(q.shape[0] + 1, np.uint64)
        np.cumsum(cnt, out=off[1:])
        out = np.zeros(max(tot, 1), np.uint32)
        RefPatricia._lib.rp_match_unique(self.h, _p(q), q.shape[0], threads, _p(cnt), _p(off), _p(out))
        return off, out[:tot]

    def close(self):
        if self.h:
            RefPatricia._lib.rp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
