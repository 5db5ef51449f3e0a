"""The C ABI library: loads, exports every declared entry point, and its
host-side pieces (generator, canonical table) behave; no GPU needed."""
import os
import re

import numpy as np
import pytest

import tagmatch_amd as tm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(tm_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    lib = tm.load()
    decl = declared_functions()
    assert len(decl) >= 30
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(tm.EXPORTED) <= decl
    assert b"gfx950" in lib.tm_version()


def test_library_has_gfx950_code_object():
    data = open(tm.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_no_device_is_reported_not_crashed():
    if tm.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = tm.load()
    assert lib.tm_dev_init(0) == tm.TM_ERR_NODEV
    assert lib.tm_mem_info(None, None) == tm.TM_ERR_STATE


def test_argument_errors_before_device_use():
    lib = tm.load()
    fib = tm.C.c_void_p()
    assert lib.tm_fib_create(tm.C.byref(fib)) == tm.TM_OK
    words = np.zeros((2, 6), np.uint32)
    words[:, 0] = 1
    off = np.array([0, 2], np.uint32)
    bad_mask = np.array([[2, 0, 0, 0, 0, 0]], np.uint32)      # not a subset of the filters
    ko = np.array([0, 1, 2], np.uint32)
    keys = np.array([5, 6], np.uint32)
    rc = lib.tm_fib_upload(fib, tm._ptr(words), 2, tm._ptr(off), tm._ptr(bad_mask), 1, tm._ptr(ko), tm._ptr(keys), None)
    assert rc == tm.TM_ERR_INVALID
    ko_bad = np.array([0, 1, 1], np.uint32)                   # filter without keys
    good_mask = np.array([[1, 0, 0, 0, 0, 0]], np.uint32)
    rc = lib.tm_fib_upload(fib, tm._ptr(words), 2, tm._ptr(off), tm._ptr(good_mask), 1, tm._ptr(ko_bad), tm._ptr(keys), None)
    assert rc == tm.TM_ERR_INVALID
    assert b"no keys" in lib.tm_last_error()
    assert lib.tm_fib_destroy(fib) == tm.TM_OK


def test_generator_is_deterministic_and_sliceable():
    g1 = tm.Generator(seed=123)
    g2 = tm.Generator(seed=123)
    a = g1.filters(5000)
    b = g2.filters(5000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # any slice is the same as the corresponding part of the whole
    w, ko, k = g1.filters(1000, first=2000)
    assert np.array_equal(w, a[0][2000:3000])
    qa, sa = g1.queries(3000, 5000)
    qb, sb = g2.queries(1000, 5000, first=1000)
    assert np.array_equal(qa[1000:2000], qb) and np.array_equal(sa[1000:2000], sb)
    assert not np.array_equal(tm.Generator(seed=124).filters(100)[0], a[0][:100])


def test_generator_statistics_match_the_workload_spec():
    g = tm.Generator()
    w, ko, k = g.filters(50000)
    q, s = g.queries(20000, 50000)
    pc = lambda a: np.unpackbits(a.view(np.uint8), axis=1).sum(1)
    nk = np.diff(ko)
    assert 1 <= nk.min() and nk.max() <= 4 and abs(nk.mean() - 2.5) < 0.05
    assert 25 < pc(w).mean() < 40                # 1-10 tags x 7 bits, collisions
    assert 85 < pc(q).mean() < 115               # 10-30 tags
    assert abs(s.mean() - 0.3) < 0.02
    trees, ifaces = k >> 16, k & 0xFFFF
    assert trees.max() == 7 and ifaces.max() == 255
    # superset queries contain their filter: at least one match each
    sup = np.nonzero(s)[0][:200]
    m = np.array([np.any(np.all((w & ~q[i]) == 0, axis=1)) for i in sup])
    assert m.all()


def test_canonicalize_merges_and_orders():
    w = np.zeros((4, 6), np.uint32)
    w[0, 0] = 2; w[1, 0] = 1; w[2, 0] = 2; w[3, 5] = 1 << 31
    ko = np.array([0, 2, 3, 5, 6], np.uint64)
    k = np.array([9, 3, 4, 3, 7, 1], np.uint32)
    cw, cko, ck = tm.canonicalize(w, ko, k)
    # bit 0 set sorts after bit 0 clear; identical filters merged, keys union
    assert cw.shape[0] == 3
    assert list(cw[:, 0]) == [0, 2, 1] and cw[0, 5] == 1 << 31
    assert list(ck[cko[1]:cko[2]]) == [3, 7, 9]


def test_cpp_api_driver_built():
    assert os.path.exists(os.path.join(ROOT, "tagmatch_amd", "lib", "tm_api_tests"))


@pytest.mark.parametrize("n_keys", [1, 31, 32, 33, 100, 2048])
def test_expand_key_bitmaps_host(n_keys):
    """tm_expand_key_bitmaps (host code, no GPU): the keys of the set bits in
    increasing order, bits at or above n_keys ignored, full words and empty
    rows, against numpy."""
    rng = np.random.default_rng(n_keys)
    kw = (n_keys + 31) // 32 + 1                        # one spare word of garbage
    nq = 3000
    rows = rng.integers(0, 2**32, size=(nq, kw), dtype=np.uint64).astype(np.uint32)
    rows[::7] = 0xFFFFFFFF                              # full rows
    rows[1::11] = 0                                     # empty rows
    rows[2::5, 0] = 0xFFFFFFFF                          # a full first word
    dictionary = np.sort(rng.choice(2**31, n_keys, replace=False)).astype(np.uint32)
    for threads in (1, 0):
        off, vals = tm.expand_key_bitmaps(rows, dictionary, threads=threads)
        bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :n_keys].astype(bool)
        assert np.array_equal(np.diff(off), bits.sum(axis=1))
        exp = np.concatenate([dictionary[b] for b in bits])
        assert np.array_equal(vals, exp)
    # capacity too small: TM_ERR_OVERFLOW, offsets still written
    lib = tm.load()
    o2 = np.zeros(nq + 1, np.uint64)
    small = np.zeros(4, np.uint32)
    rc = lib.tm_expand_key_bitmaps(tm._ptr(rows), nq, kw, tm._ptr(dictionary), n_keys, tm._ptr(o2), tm._ptr(small),
                                   small.size, 2)
    assert rc == tm.TM_ERR_OVERFLOW and np.array_equal(o2, off)


@pytest.mark.parametrize("n_keys", [1, 33, 2048])
def test_expand_key_rows_packed_host(n_keys):
    """tm_expand_key_rows_packed (host code, no GPU): TM_ROW_FULL entries
    give the whole dictionary, the others their packed row; equal to the
    expansion of the unpacked rows; an index past the packed rows is
    rejected."""
    rng = np.random.default_rng(100 + n_keys)
    kw = (n_keys + 31) // 32
    nq = 2000
    full = rng.random(nq) < 0.6
    n_rows = int((~full).sum())
    packed = rng.integers(0, 2**32, size=(n_rows, kw), dtype=np.uint64).astype(np.uint32)
    if n_keys % 32:
        packed[:, -1] &= np.uint32((1 << (n_keys % 32)) - 1)
    row_of = np.full(nq, tm.ROW_FULL, np.uint32)
    row_of[~full] = np.arange(n_rows, dtype=np.uint32)
    rows = np.zeros((nq, kw), np.uint32)
    rows[full] = 0xFFFFFFFF
    if n_keys % 32:
        rows[full, -1] = (1 << (n_keys % 32)) - 1
    rows[~full] = packed
    dictionary = np.sort(rng.choice(2**31, n_keys, replace=False)).astype(np.uint32)
    eo, ev = tm.expand_key_bitmaps(rows, dictionary)
    for threads in (1, 0):
        off, vals = tm.expand_key_rows_packed(row_of, packed, dictionary, threads=threads)
        assert np.array_equal(off, eo) and np.array_equal(vals, ev)
    assert np.all(np.diff(off)[full] == n_keys)
    bad = row_of.copy()
    bad[np.flatnonzero(~full)[0]] = n_rows
    with pytest.raises(tm.TagMatchError):
        tm.expand_key_rows_packed(bad, packed, dictionary)
