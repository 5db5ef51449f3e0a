"""Reference file formats (tm_io_*): filter tables and queries in the
reference's binary, ASCII and partitioned forms.  Host only (no GPU).

Pinned by the filter encoding cases of the reference's test/test_io.cc
(tests/golden/io_vectors.json), by records packed by hand from the format
(reference fib.cc:14-35, fib.hh:41-56, query.cc:11-40), and by round trips."""
import json
import os
import struct

import numpy as np
import pytest

import tagmatch_amd as tm

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "io_vectors.json")


def test_reference_encoding_cases(tmp_path):
    """ASCII filter -> words -> 24 binary bytes, as test_io.cc expects."""
    cases = json.load(open(GOLDEN))["cases"]
    assert len(cases) == 5
    for c in cases:
        pa = tmp_path / f"{c['name']}.txt"
        pa.write_text(f"! {c['ascii']}\n")
        q = tm.read_query_file(str(pa), tm.IO_ASCII)
        assert q.shape == (1, 6)
        pb = tmp_path / f"{c['name']}.bin"
        tm.write_query_file(str(pb), q, tm.IO_BINARY)
        assert pb.read_bytes().hex() == c["binary_hex"], c["name"]
        # and back: binary -> ASCII gives the same 192 characters
        q2 = tm.read_query_file(str(pb), tm.IO_BINARY)
        tm.write_query_file(str(pa), q2, tm.IO_ASCII)
        assert pa.read_text() == f"! {c['ascii']}\n"


def test_fib_records_packed_by_hand(tmp_path):
    """Binary FIB entries: filter words little-endian, u32 count, u32 keys;
    the partitioned form puts a u32 partition id in front."""
    w = np.array([[1, 0, 0, 0, 0, 0x80000000], [0xDEADBEEF, 2, 3, 4, 5, 6]], np.uint32)
    keys = [[7, 9, 65536 * 3 + 12], []]
    plain = b"".join(struct.pack("<6I", *w[i]) + struct.pack("<I", len(keys[i])) +
                     struct.pack(f"<{len(keys[i])}I", *keys[i]) for i in range(2))
    (tmp_path / "a.bin").write_bytes(plain)
    rw, ko, kv, parts = tm.read_fib_file(str(tmp_path / "a.bin"), tm.IO_BINARY)
    assert np.array_equal(rw, w)
    assert ko.tolist() == [0, 3, 3]
    assert kv.tolist() == keys[0]
    part = b"".join(struct.pack("<I", 40 + i) + struct.pack("<6I", *w[i]) + struct.pack("<I", len(keys[i])) +
                    struct.pack(f"<{len(keys[i])}I", *keys[i]) for i in range(2))
    (tmp_path / "p.bin").write_bytes(part)
    rw, ko, kv, parts = tm.read_fib_file(str(tmp_path / "p.bin"), tm.IO_BINARY_PARTITIONED)
    assert np.array_equal(rw, w) and parts.tolist() == [40, 41] and kv.tolist() == keys[0]
    # writing gives the same bytes
    tm.write_fib_file(str(tmp_path / "b.bin"), w, ko, kv, tm.IO_BINARY)
    assert (tmp_path / "b.bin").read_bytes() == plain
    tm.write_fib_file(str(tmp_path / "q.bin"), w, ko, kv, tm.IO_BINARY_PARTITIONED, partitions=[40, 41])
    assert (tmp_path / "q.bin").read_bytes() == part


def test_ascii_forms(tmp_path):
    """ASCII entries: '+ <bits> keys', other lines skipped; a filter string
    shorter than 192 bits leaves the remaining bits zero."""
    bits = "1" + "0" * 30 + "1" + "0" * 159 + "1"    # bits 0, 31 and 191
    text = "# a comment line\n" f"+ {bits} 5 6 7\n" "\n" "+ 0101 42\n" "! 1\n"
    (tmp_path / "f.txt").write_text(text)
    w, ko, kv, _ = tm.read_fib_file(str(tmp_path / "f.txt"), tm.IO_ASCII)
    assert w.shape == (2, 6)
    assert w[0].tolist() == [0x80000001, 0, 0, 0, 0, 0x80000000]
    assert w[1].tolist() == [0b1010, 0, 0, 0, 0, 0]
    assert ko.tolist() == [0, 3, 4] and kv.tolist() == [5, 6, 7, 42]
    q = tm.read_query_file(str(tmp_path / "f.txt"), tm.IO_ASCII)
    assert q.tolist() == [[1, 0, 0, 0, 0, 0]]


@pytest.mark.parametrize("fmt", [tm.IO_BINARY, tm.IO_ASCII, tm.IO_BINARY_PARTITIONED])
def test_round_trip(tmp_path, fmt):
    rng = np.random.default_rng(77 + fmt)
    n = 500
    w = rng.integers(0, 2**32, size=(n, 6), dtype=np.uint64).astype(np.uint32)
    counts = rng.integers(0, 9, n)
    ko = np.zeros(n + 1, np.uint32)
    np.cumsum(counts, out=ko[1:])
    kv = rng.integers(0, 2**32, int(ko[-1]), dtype=np.uint64).astype(np.uint32)
    parts = rng.integers(0, 1000, n).astype(np.uint32)
    p = str(tmp_path / "t")
    tm.write_fib_file(p, w, ko, kv, fmt, partitions=parts)
    rw, rko, rkv, rparts = tm.read_fib_file(p, fmt)
    assert np.array_equal(rw, w) and np.array_equal(rko, ko) and np.array_equal(rkv, kv)
    if fmt == tm.IO_BINARY_PARTITIONED:
        assert np.array_equal(rparts, parts)
    if fmt != tm.IO_BINARY_PARTITIONED:
        tm.write_query_file(p + ".q", w, fmt)
        assert np.array_equal(tm.read_query_file(p + ".q", fmt), w)


def test_errors(tmp_path):
    w = np.ones((2, 6), np.uint32)
    tm.write_fib_file(str(tmp_path / "a.bin"), w, [0, 2, 3], [1, 2, 3])
    data = (tmp_path / "a.bin").read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[:-2])               # last key cut short
    with pytest.raises(tm.TagMatchError) as e:
        tm.read_fib_file(str(tmp_path / "cut.bin"))
    assert e.value.code == tm.TM_ERR_INVALID and "truncated" in str(e.value)
    (tmp_path / "q.bin").write_bytes(b"\0" * 30)                 # not a multiple of 24 bytes
    with pytest.raises(tm.TagMatchError):
        tm.read_query_file(str(tmp_path / "q.bin"))
    with pytest.raises(tm.TagMatchError):
        tm.read_fib_file(str(tmp_path / "missing.bin"))
    with pytest.raises(tm.TagMatchError):
        tm.read_query_file(str(tmp_path / "q.bin"), tm.IO_BINARY_PARTITIONED)   # no partitioned queries
