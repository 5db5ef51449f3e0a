"""The workload tools (tm_io_convert, tm_io_sample, tm_io_split) against the
reference's own convert_workload, sample_workload and split_workload
programs, built from their sources into oracle/_ref (skipped where they are
absent), on generated files; and the properties that hold without them.
Host code only: no GPU."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

REF = orc.REF_DIR


def ref_tool(name):
    if not orc.ref_available(name):
        pytest.skip(f"oracle/_ref/{name} not built (reference sources absent)")
    return os.path.join(REF, name)


@pytest.fixture(scope="module")
def table():
    cw, cko, ck, q, _ = tm.workload(3000, 400, seed=77)
    return cw, cko.astype(np.uint32), ck, q


def write_fib(path, table, fmt=tm.IO_BINARY, partitions=None):
    cw, cko, ck, q = table
    tm.write_fib_file(str(path), cw, cko, ck, fmt=fmt, partitions=partitions)


def read_bytes(p):
    with open(p, "rb") as f:
        return f.read()


# ------------------------------------------------------------------ convert

def test_convert_fib_binary_to_ascii_as_reference(table, tmp_path):
    exe = ref_tool("convert_workload")
    src = tmp_path / "f.bin"
    write_fib(src, table)
    n = tm.convert_file(str(src), tm.IO_BINARY, str(tmp_path / "ours.txt"), tm.IO_ASCII, tm.KIND_FIB)
    assert n == len(table[0])
    subprocess.run([exe, f"in={src}", f"out={tmp_path / 'ref.txt'}", "-a", "-F"], check=True)
    assert read_bytes(tmp_path / "ours.txt") == read_bytes(tmp_path / "ref.txt")
    # and back: the reference's ASCII FIB reader takes the first key from the
    # file stream, not the line (fib.cc:47-69), so its output is not usable;
    # ours must reproduce the original binary file
    tm.convert_file(str(tmp_path / "ours.txt"), tm.IO_ASCII, str(tmp_path / "back.bin"), tm.IO_BINARY, tm.KIND_FIB)
    assert read_bytes(tmp_path / "back.bin") == read_bytes(src)


def test_convert_partitioned_fib_and_queries_as_reference(table, tmp_path):
    exe = ref_tool("convert_workload")
    cw, cko, ck, q = table
    parts = (np.arange(len(cw)) // 37).astype(np.uint32)
    src = tmp_path / "pf.bin"
    write_fib(src, table, fmt=tm.IO_BINARY_PARTITIONED, partitions=parts)
    tm.write_query_file(str(tmp_path / "q.bin"), q)
    cases = [("pf.bin", tm.IO_BINARY_PARTITIONED, tm.IO_ASCII_PARTITIONED, tm.KIND_FIB, ["-a", "-f"]),
             ("q.bin", tm.IO_BINARY, tm.IO_ASCII, tm.KIND_QUERIES, ["-a", "-q"])]
    for name, fi, fo, kind, flags in cases:
        tm.convert_file(str(tmp_path / name), fi, str(tmp_path / (name + ".ours")), fo, kind)
        subprocess.run([exe, f"in={tmp_path / name}", f"out={tmp_path / (name + '.ref')}", *flags], check=True)
        assert read_bytes(tmp_path / (name + ".ours")) == read_bytes(tmp_path / (name + ".ref")), name
        # ASCII back to binary, ours and the reference's
        tm.convert_file(str(tmp_path / (name + ".ours")), fo, str(tmp_path / (name + ".b.ours")), fi, kind)
        subprocess.run([exe, f"in={tmp_path / (name + '.ours')}", f"out={tmp_path / (name + '.b.ref')}",
                        flags[1]], check=True)
        assert read_bytes(tmp_path / (name + ".b.ours")) == read_bytes(tmp_path / name), name
        assert read_bytes(tmp_path / (name + ".b.ref")) == read_bytes(tmp_path / name), name


def test_convert_between_partitioned_and_plain(table, tmp_path):
    cw, cko, ck, q = table
    write_fib(tmp_path / "f.bin", table)
    tm.convert_file(str(tmp_path / "f.bin"), tm.IO_BINARY, str(tmp_path / "p.txt"), tm.IO_ASCII_PARTITIONED,
                    tm.KIND_FIB)
    w, ko, k, parts = tm.read_fib_file(str(tmp_path / "p.txt"), fmt=tm.IO_ASCII_PARTITIONED)
    assert np.array_equal(w, cw) and np.array_equal(ko, cko) and np.array_equal(k, ck)
    assert not parts.any()
    with pytest.raises(tm.TagMatchError):
        tm.convert_file(str(tmp_path / "f.bin"), tm.IO_BINARY, str(tmp_path / "x"), tm.IO_ASCII_PARTITIONED,
                        tm.KIND_QUERIES)


# ------------------------------------------------------------------ sample

@pytest.mark.parametrize("n_total,k,seed", [(3000, 100, 1), (1000, 999, 7), (2500, 2500, 3), (500, 1, 99),
                                            (5000, 300, 12345)])
def test_sample_fib_as_reference(table, tmp_path, n_total, k, seed):
    exe = ref_tool("sample_workload")
    src = tmp_path / "f.bin"
    write_fib(src, table)
    n = tm.sample_file(str(src), tm.IO_BINARY, str(tmp_path / "ours.bin"), tm.KIND_FIB, n_total, k, seed)
    subprocess.run([exe, f"in={src}", f"out={tmp_path / 'ref.bin'}", f"N={n_total}", f"K={k}", f"seed={seed}"],
                   check=True)
    assert read_bytes(tmp_path / "ours.bin") == read_bytes(tmp_path / "ref.bin")
    # exactly k when the file holds the n_total records sampled from
    assert n == k if n_total <= len(table[0]) else n <= k


@pytest.mark.parametrize("fmt", [tm.IO_BINARY, tm.IO_ASCII])
def test_sample_queries_as_reference(table, tmp_path, fmt):
    exe = ref_tool("sample_workload")
    q = table[3]
    src = tmp_path / "q"
    tm.write_query_file(str(src), q, fmt=fmt)
    flags = ["-p"] + (["-a"] if fmt == tm.IO_ASCII else [])
    for n_total, k, seed in ((400, 50, 5), (300, 299, 8), (400, 400, 1)):
        n = tm.sample_file(str(src), fmt, str(tmp_path / "ours"), tm.KIND_QUERIES, n_total, k, seed)
        subprocess.run([exe, f"in={src}", f"out={tmp_path / 'ref'}", f"N={n_total}", f"K={k}", f"seed={seed}",
                        *flags], check=True)
        assert read_bytes(tmp_path / "ours") == read_bytes(tmp_path / "ref"), (n_total, k, seed)
        assert n <= k


def test_sample_is_an_ordered_subset_of_uniform_frequency(table, tmp_path):
    """Without the reference: with records to spare the output is an
    in-order subsequence of exactly K records, nearly all from the first N,
    and over many seeds each of the first N is kept about K/N of the time."""
    q = np.zeros((100, 6), np.uint32)
    q[:, 0] = np.arange(100)                    # record i holds i
    tm.write_query_file(str(tmp_path / "q.bin"), q)
    hits = np.zeros(100)
    # seeds spread over the generator's range (minstd's first draws from
    # small seeds are small, as in the reference tool)
    seeds = np.random.default_rng(1).integers(1, 2**31 - 1, 400)
    for seed in seeds:
        tm.sample_file(str(tmp_path / "q.bin"), tm.IO_BINARY, str(tmp_path / "s.bin"), tm.KIND_QUERIES, 80, 20,
                       int(seed))
        idx = tm.read_query_file(str(tmp_path / "s.bin"))[:, 0].astype(np.int64)
        assert len(idx) == 20 and np.all(np.diff(idx) > 0)
        hits[idx] += 1
    freq = hits[:80] / 400
    assert freq.min() > 0.15 and freq.max() < 0.35 and abs(freq.mean() - 0.25) < 0.01
    assert hits[80:].sum() < 0.02 * hits.sum()


# ------------------------------------------------------------------ split

@pytest.mark.parametrize("m,k", [(100, 2), (37, 0), (500, 5), (1, 2), (4000, 1)])
@pytest.mark.parametrize("order", ["ascending", "generated"])
def test_split_as_reference(table, tmp_path, m, k, order):
    exe = ref_tool("split_workload")
    cw, cko, ck, q = table
    if order == "ascending":       # the canonical table is ascending: kept as it is
        w, ko, kv = cw, cko, ck
    else:                          # any other order is sorted descending
        perm = np.random.default_rng(5).permutation(len(cw))
        cnt = np.diff(cko.astype(np.int64))[perm]
        ko = np.zeros(len(cw) + 1, np.uint32)
        np.cumsum(cnt, out=ko[1:])
        kv = np.concatenate([ck[cko[i]:cko[i + 1]] for i in perm])
        w = cw[perm]
    src = tmp_path / "f.bin"
    tm.write_fib_file(str(src), w, ko, kv)
    n = tm.split_file(str(src), tm.IO_BINARY, str(tmp_path / "p.ours"), str(tmp_path / "f.ours"), m, k)
    subprocess.run([exe, f"in={src}", f"p={tmp_path / 'p.ref'}", f"f={tmp_path / 'f.ref'}", f"m={m}", f"k={k}"],
                   check=True)
    assert read_bytes(tmp_path / "p.ours") == read_bytes(tmp_path / "p.ref")
    assert read_bytes(tmp_path / "f.ours") == read_bytes(tmp_path / "f.ref")
    prefixes = tm.read_prefix_file(str(tmp_path / "p.ours"))
    assert len(prefixes["size"]) == n and int(prefixes["size"].sum()) == len(cw)
    assert np.array_equal(prefixes["partition"], np.arange(n))
    assert (prefixes["size"] <= max(m, 1)).all()
    # ASCII input gives the ASCII form of the same outputs (the reference
    # cannot read ASCII FIB entries, fib.cc:47-69, so this compares with our
    # binary outputs converted)
    tm.convert_file(str(src), tm.IO_BINARY, str(tmp_path / "f.txt"), tm.IO_ASCII, tm.KIND_FIB)
    tm.split_file(str(tmp_path / "f.txt"), tm.IO_ASCII, str(tmp_path / "p.txt"), str(tmp_path / "f2.txt"), m, k)
    tm.convert_file(str(tmp_path / "p.ours"), tm.IO_BINARY, str(tmp_path / "p2.txt"), tm.IO_ASCII, tm.KIND_PREFIXES)
    assert read_bytes(tmp_path / "p.txt") == read_bytes(tmp_path / "p2.txt")
    tm.convert_file(str(tmp_path / "f.ours"), tm.IO_BINARY_PARTITIONED, str(tmp_path / "f3.txt"),
                    tm.IO_ASCII_PARTITIONED, tm.KIND_FIB)
    assert read_bytes(tmp_path / "f2.txt") == read_bytes(tmp_path / "f3.txt")


def test_prefix_files_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 50
    words = rng.integers(0, 2**32, size=(n, 6), dtype=np.uint64).astype(np.uint32)
    length = rng.integers(0, 193, n).astype(np.uint32)
    # the ASCII form holds only the first `length` bits
    for i in range(n):
        for b in range(int(length[i]), 192):
            words[i, b // 32] &= ~np.uint32(1 << (b % 32))
    part = np.arange(n, dtype=np.uint32)
    size = rng.integers(1, 10**6, n).astype(np.uint32)
    for fmt in (tm.IO_BINARY, tm.IO_ASCII):
        tm.write_prefix_file(str(tmp_path / "p"), words, length, part, size, fmt=fmt)
        r = tm.read_prefix_file(str(tmp_path / "p"), fmt=fmt)
        assert np.array_equal(r["words"], words) and np.array_equal(r["length"], length)
        assert np.array_equal(r["partition"], part) and np.array_equal(r["size"], size)
    assert os.path.getsize(tmp_path / "p") > 0
    tm.write_prefix_file(str(tmp_path / "b"), words, length, part, size)
    assert os.path.getsize(tmp_path / "b") == 33 * n
