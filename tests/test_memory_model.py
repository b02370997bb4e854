"""The cross-workgroup hand-offs (grid reductions, flags, counters) of the
cooperative kernels publish with an agent-scope release: an L2 write-back
(buffer_wbl2) that must have completed before the flag or counter update a
reader on another XCD acquires.  hipcc for gfx950 drops that wait when the
source already waited on vmcnt before the release, so the kernels release
through rnpx::release_agent (fence + explicit wait).  This compiles the
kernel sources to gfx950 assembly (CPU only, with the library's flags) and
checks that every buffer_wbl2 is followed by s_waitcnt vmcnt(0) before the
next memory instruction."""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rnp2_amd", "csrc")
SOURCES = ["rnpx_trsm.hip", "rnpx_cqr.hip", "rnpx_lu.hip", "rnpx_panel.hip", "rnpx_batched.hip"]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def unwaited_writebacks(asm):
    """Number of buffer_wbl2 not followed by s_waitcnt vmcnt(0) before the
    next vector-memory instruction, and the total."""
    total = bad = 0
    pending = False
    for line in asm.splitlines():
        s = line.strip()
        if s.startswith("buffer_wbl2"):
            total += 1
            pending = True
            continue
        if pending and s.startswith("s_waitcnt") and re.search(r"vmcnt\(0\)", s):
            pending = False
            continue
        if pending and re.match(r"(global_|buffer_|flat_|scratch_)", s):
            bad += 1
            pending = False
    return bad, total


def compile_asm(src, out_dir):
    out = os.path.join(out_dir, src + ".s")
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "--cuda-device-only", "-S", "-O3", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return open(out).read()


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_every_l2_writeback_is_waited_on():
    with tempfile.TemporaryDirectory() as d:
        with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            asm = dict(zip(SOURCES, ex.map(lambda s: compile_asm(s, d), SOURCES)))
    report = {s: unwaited_writebacks(a) for s, a in asm.items()}
    print(report)
    assert sum(t for _, t in report.values()) > 0, "no release found: the check would pass vacuously"
    assert all(b == 0 for b, _ in report.values()), report


def test_checker_flags_an_unwaited_writeback():
    bad_asm = "\tglobal_store_dwordx2 v1, v[2:3], s[4:5]\n\ts_waitcnt vmcnt(0)\n\tbuffer_wbl2 sc1\n" \
              "\tglobal_store_dword v0, v1, s[6:7] sc1\n"
    good_asm = "\tglobal_store_dwordx2 v1, v[2:3], s[4:5]\n\tbuffer_wbl2 sc1\n\ts_waitcnt vmcnt(0)\n" \
               "\tglobal_store_dword v0, v1, s[6:7] sc1\n"
    assert unwaited_writebacks(bad_asm) == (1, 1)
    assert unwaited_writebacks(good_asm) == (0, 1)
