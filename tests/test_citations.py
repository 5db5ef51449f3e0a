"""The reference line ranges cited in this repository's sources and notes
exist, and the ranges that name a function hold it.  Runs only where the
reference sources are mounted (skipped elsewhere)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("REF", "/root/reference")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not mounted")

SCANNED = ["include", "oracle", "tagmatch_amd", "tests", "bench.py", "__graft_entry__.py", "DESIGN.md",
           "INTEGRATION.md", "README.md"]
CITE = re.compile(r"([A-Za-z_][A-Za-z0-9_]*\.(?:cc|hh|cu|h))((?::\d+(?:-\d+)?)(?:,\s*:\d+(?:-\d+)?)*)")

# (file, first, last, text that must appear in that range)
NAMED = [
    ("partitioning.cc", 230, 250, "nonzero_mask_partitioning"),
    ("partitioning.cc", 262, 278, "balancing_pivot"),
    ("partitioning.cc", 280, 340, "do_balanced_partitioning"),
    ("partitioning.cc", 361, 411, "balanced_partitioning"),
    ("partitioning.cc", 92, 113, "compute_frequencies"),
    ("partitioning.cc", 133, 149, "split_zero"),
    ("partitioning.cc", 389, 402, "partition.filter &="),
    ("back_end.cc", 363, 453, "compile_fibs"),
    ("back_end.cc", 243, 286, "process_batch"),
    ("back_end.cc", 288, 301, "add_output"),
    ("back_end.cc", 131, 311, "struct stream_handle"),
    ("back_end.cc", 69, 88, "tmp_fib_users"),
    ("back_end.cc", 490, 516, "CPU_BACKEND"),
    ("tagmatch.cc", 98, 165, "tagmatch::consolidate"),
    ("tagmatch.cc", 36, 60, "keyset_change::apply"),
    ("tagmatch.cc", 200, 204, "tagmatch::match"),
    ("tagmatch.cc", 206, 210, "tagmatch::match_unique"),
    ("front_end.cc", 465, 477, "do_match"),
    ("front_end.cc", 249, 267, "first_pending"),
    ("front_end.cc", 350, 376, "partition_queue::enqueue"),
    ("front_end.cc", 421, 461, "front_end::add_partition"),
    ("front_end.cc", 542, 551, "front_end::match"),
    ("front_end.cc", 563, 598, "match_loop"),
    ("front_end.cc", 693, 771, "front_end::stop"),
    ("gpu.cu", 878, 912, "gpu::run_kernel"),
    ("tagmatch_query.hh", 99, 119, "finalize_matching"),
    ("tagmatch_query.hh", 105, 109, "match_unique"),
    ("test_matcher.cc", 74, 81, "match"),
    ("bitvector.hh", 290, 296, "subset_of"),
    ("bitvector.hh", 389, 397, "operator <"),
    ("bitvector.hh", 412, 418, "leftmost_diff"),
    ("fib.cc", 14, 35, "write_binary"),
    ("query.cc", 11, 19, "write_binary"),
    ("convert_workload.cc", 12, 59, "convert_filters"),
    ("split_workload.cc", 37, 119, "split_on_prefix"),
]


def ref_files():
    out = {}
    for d, _, fs in os.walk(REF):
        for f in fs:
            out.setdefault(f, os.path.join(d, f))
    return out


def lines_of(path):
    with open(path, errors="replace") as f:
        return f.read().split("\n")


def test_cited_ranges_exist():
    files = ref_files()
    own = {f for d, _, fs in os.walk(ROOT) for f in fs}
    bad = []
    for r in SCANNED:
        p0 = os.path.join(ROOT, r)
        paths = [p0] if os.path.isfile(p0) else [os.path.join(d, f) for d, _, fs in os.walk(p0) for f in fs]
        for p in paths:
            if "/lib/" in p or "/_ref" in p or "__pycache__" in p or p.endswith((".so", ".npz", ".pyc", ".json")):
                continue
            text = open(p, errors="replace").read()
            for ln, line in enumerate(text.split("\n"), 1):
                for m in CITE.finditer(line):
                    name = m.group(1)
                    # file names this repository also uses for its own files
                    # are cited with "reference" in front or not at all
                    if name not in files or (name in own and "reference" not in line):
                        continue
                    n = len(lines_of(files[name]))
                    for a, b in re.findall(r":(\d+)(?:-(\d+))?", m.group(2)):
                        last = int(b or a)
                        if int(a) < 1 or last > n or last < int(a):
                            bad.append(f"{os.path.relpath(p, ROOT)}:{ln}: {name}:{a}-{last} ({n} lines)")
    assert not bad, "citations outside their files:\n" + "\n".join(bad)


@pytest.mark.parametrize("name,a,b,text", NAMED)
def test_named_ranges_hold_their_function(name, a, b, text):
    files = ref_files()
    assert name in files, name
    body = "\n".join(lines_of(files[name])[a - 1:b])
    assert text in body, f"{name}:{a}-{b} does not contain {text!r}"
