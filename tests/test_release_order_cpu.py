"""CPU checks of the release order of the persistent solvers' grid sums.

A workgroup publishes a partial-sum record only after the vectors it wrote
for the other workgroups are visible to them.  At agent scope the release
fence writes back the L2 of the writer's XCD (buffer_wbl2), and the record
stores must wait for that write-back (s_waitcnt vmcnt(0)).  The compiler
drops that wait after a bare fence when no other vector memory operation is
outstanding; a record then occasionally reached another XCD before the data
it published, and a solve read stale neighbour values (trajectories of an
instance in a batch differed from its lone run by ~1e-11).

* every release in the device sources goes through release_for_records
  (fence + explicit wait), the atomics after it relaxed: a release atomic
  gets the same write-back with its wait dropped (seen on the grid barrier's
  arrival increment);
* in the gfx950 assembly of every device source, each buffer_wbl2 is
  followed by a vmcnt(0) wait before the next vector store or atomic.
"""
import concurrent.futures
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lf_dem_amd", "csrc")
DEVICE_SOURCES = ["lfd_kernels.hip", "lfd_agg.hip", "lfd_agg_large.hip", "lfd_noalloc.hip"]


def test_releases_go_through_the_helper():
    hits = []
    for name in sorted(os.listdir(CSRC)):
        if not name.endswith((".hip", ".hpp", ".cpp")):
            continue
        for k, line in enumerate(open(os.path.join(CSRC, name)), 1):
            if "__atomic_signal_fence" in line:       # compiler-only barrier, no instruction
                continue
            if re.search(r"__ATOMIC_(RELEASE|ACQ_REL|SEQ_CST)", line):
                hits.append((name, k))
    assert len(hits) == 1 and hits[0][0] == "lfd_device.hpp", hits
    text = open(os.path.join(CSRC, "lfd_device.hpp")).read()
    body = re.search(r"void release_for_records\(\)\s*\{(.*?)\n\}", text, re.S).group(1)
    assert "__ATOMIC_RELEASE" in body and "drain_vmem()" in body


def _asm(src, out_dir):
    from lf_dem_amd import build
    out = os.path.join(out_dir, os.path.splitext(src)[0] + ".s")
    cmd = [build.hipcc()] + build.flags() + ["--cuda-device-only", "-S", "-x", "hip",
                                             os.path.join(CSRC, src), "-o", out]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return out


def _unwaited_writebacks(lines):
    """Line numbers of buffer_wbl2 with no vmcnt(0) wait before the next
    vector store or atomic (straight-line scan)."""
    bad = []
    for i, line in enumerate(lines):
        if "buffer_wbl2" not in line:
            continue
        waited = False
        for nxt in lines[i+1:i+600]:
            s = nxt.strip()
            if s.startswith("s_waitcnt") and "vmcnt(0)" in s:
                waited = True
            if re.match(r"(global|buffer|flat)_(store|atomic)", s):
                break
        if not waited:
            bad.append(i+1)
    return bad


def test_scan_flags_a_writeback_without_wait():
    lines = ["buffer_wbl2 sc1", "s_waitcnt lgkmcnt(0)", "global_store_dwordx2 v1, v[2:3], s[4:5] sc1",
             "buffer_wbl2 sc1", "s_waitcnt vmcnt(0)", "global_store_dwordx2 v1, v[2:3], s[4:5] sc1"]
    assert _unwaited_writebacks(lines) == [1]


def test_every_writeback_is_waited_for_in_the_gfx950_assembly(tmp_path):
    try:
        from lf_dem_amd import build
        build.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    with concurrent.futures.ThreadPoolExecutor(4) as ex:
        outs = list(ex.map(lambda s: _asm(s, str(tmp_path)), DEVICE_SOURCES))
    total = 0
    for src, out in zip(DEVICE_SOURCES, outs):
        lines = open(out).read().split("\n")
        total += sum("buffer_wbl2" in l for l in lines)
        assert _unwaited_writebacks(lines) == [], src
    assert total > 0
