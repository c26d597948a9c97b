"""Audit of the compiled gfx950 code (CPU only: hipcc cross-compiles).

Kernels that copy operands with LDS-DMA loads (`global_load_lds_*`) must not
touch scratch memory, and no NT kernel may write memory with AGPR data.  On
some MI355X boards the forms that did (accumulators stored straight from
AGPRs; accumulators spilled to scratch) returned wrong words while the same
products in register-only form were exact, so the rule is checked on the
code the compiler actually emits, not on the source.

Two orderings the LDS-DMA main loop relies on are pinned the same way:
  * every s_barrier in an LDS-DMA kernel is reached through an
    `s_waitcnt vmcnt(0)` with no LDS-DMA copy issued in between, in the same
    straight-line block: the copies count on vmcnt, and a wave may only pass
    the barrier once its own copies have landed in LDS, so that the other
    waves read complete stages after it;
  * no LDS-DMA copy directly follows an instruction that writes M0 (the copy
    takes its LDS base from M0, and an SALU write of M0 needs one wait state
    before an LDS-DMA instruction reads it).
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

from ferray_amd import build as fb

HIPCC = fb.HIPCC if os.path.exists(fb.HIPCC) else shutil.which("hipcc")
pytestmark = pytest.mark.skipif(HIPCC is None, reason="hipcc not available")

# Sources whose kernels use the LDS-DMA main loop (mfma_nt.hpp).
DMA_SOURCES = ["gemm_nt.hip", "ozaki.hip", "sgemm_split.hip", "gemm_real.hip"]

AGPR_WRITE = re.compile(r"^\s*(global|buffer|flat|scratch)_(store|atomic)\w*\s.*,\s*a[\[0-9]")
LDS_DMA = re.compile(r"^\s*(global|buffer)_load_lds_\w+")
M0_WRITE = re.compile(r"^\s*s_\w+\s+m0\b")


def _kernels(asm: str) -> dict[str, list[str]]:
    """Instruction lines of each function in a device-only .s file."""
    out: dict[str, list[str]] = {}
    cur = None
    for line in asm.splitlines():
        m = re.match(r"^(_Z\S+):", line)
        if m:
            cur = m.group(1)
            out[cur] = []
            continue
        if cur is None:
            continue
        if "s_endpgm" in line or line.startswith("\t.size\t") or line.startswith(".Lfunc_end"):
            cur = None
            continue
        out[cur].append(line)
    return out


def _instructions(body: list[str]) -> list[str]:
    """Instruction and label lines, comments and directives dropped."""
    out = []
    for line in body:
        code = line.split(";", 1)[0].rstrip()
        if not code.strip() or code.lstrip().startswith("."):
            if re.match(r"^\.LBB\w*:", code):
                out.append(code)          # a label: another path may enter here
            continue
        out.append(code)
    return out


def barrier_violations(body: list[str]) -> list[int]:
    """Indexes of s_barrier lines not reached through `s_waitcnt vmcnt(0)`
    in the same straight-line block with no LDS-DMA copy after it."""
    ins = _instructions(body)
    bad = []
    for i, line in enumerate(ins):
        if not line.strip().startswith("s_barrier"):
            continue
        ok = False
        for j in range(i - 1, -1, -1):
            prev = ins[j].strip()
            if prev.endswith(":") or prev.startswith("s_cbranch") or prev.startswith("s_branch"):
                break                     # block boundary before any wait
            if LDS_DMA.match(ins[j]):
                break                     # a copy issued after the last wait
            if prev.startswith("s_waitcnt") and "vmcnt(0)" in prev:
                ok = True
                break
        if not ok:
            bad.append(i)
    return bad


def m0_violations(body: list[str]) -> list[int]:
    """Indexes of LDS-DMA copies whose preceding instruction writes M0."""
    ins = [l for l in _instructions(body) if not l.strip().endswith(":")]
    return [i for i in range(1, len(ins)) if LDS_DMA.match(ins[i]) and M0_WRITE.match(ins[i - 1])]


@pytest.fixture(scope="module")
def device_asm(tmp_path_factory):
    d = tmp_path_factory.mktemp("asm")
    flags = [f for f in fb.CFLAGS if not f.startswith("-W")]
    asm = {}
    for src in DMA_SOURCES:
        out = d / (src + ".s")
        cmd = [HIPCC, *flags, "--cuda-device-only", "-S", "-c",
               os.path.join(fb.CSRC, src), "-o", str(out)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-4000:]
        asm[src] = _kernels(out.read_text())
    return asm


@pytest.mark.parametrize("src", DMA_SOURCES)
def test_lds_dma_kernels_use_no_scratch(device_asm, src):
    dma = {k: v for k, v in device_asm[src].items()
           if any("global_load_lds" in l for l in v)}
    assert dma, f"no LDS-DMA kernel found in {src}"
    for name, body in dma.items():
        bad = [l.strip() for l in body if "scratch_" in l]
        assert not bad, f"{name}: {len(bad)} scratch accesses, first {bad[0]}"


@pytest.mark.parametrize("src", DMA_SOURCES)
def test_no_memory_write_with_agpr_data(device_asm, src):
    for name, body in device_asm[src].items():
        bad = [l.strip() for l in body if AGPR_WRITE.match(l)]
        assert not bad, f"{name}: {len(bad)} writes with AGPR data, first {bad[0]}"


@pytest.mark.parametrize("src", DMA_SOURCES)
def test_lds_dma_barriers_wait_for_the_copies(device_asm, src):
    dma = {k: v for k, v in device_asm[src].items()
           if any(LDS_DMA.match(l) for l in v)}
    assert dma, f"no LDS-DMA kernel found in {src}"
    for name, body in dma.items():
        ins = _instructions(body)
        assert any(l.strip().startswith("s_barrier") for l in ins), name
        bad = barrier_violations(body)
        assert not bad, f"{name}: s_barrier without vmcnt(0) before it: {ins[bad[0]-3:bad[0]+1]}"


@pytest.mark.parametrize("src", DMA_SOURCES)
def test_lds_dma_copies_do_not_follow_an_m0_write(device_asm, src):
    for name, body in device_asm[src].items():
        assert not m0_violations(body), name


def test_audit_sees_a_planted_violation():
    """The parser flags both forms on a hand-written listing."""
    listing = "\n".join([
        "_Zfake:", "\tglobal_load_lds_dwordx4 v[0:1], off",
        "\tscratch_store_dwordx4 off, a[0:3], off", "\tglobal_store_dword v[2:3], a7, off",
        "\tglobal_store_dword v[2:3], v7, off", "\ts_endpgm"])
    body = _kernels(listing)["_Zfake"]
    assert sum("scratch_" in l for l in body) == 1
    assert sum(bool(AGPR_WRITE.match(l)) for l in body) == 2


def test_ordering_audit_sees_planted_violations():
    good = ["\ts_mov_b32 m0, s4", "\ts_nop 0", "\tglobal_load_lds_dwordx4 v[0:1], off",
            "\ts_waitcnt vmcnt(0) lgkmcnt(0)", "\ts_barrier"]
    assert barrier_violations(good) == [] and m0_violations(good) == []
    # the copy issued after the wait
    late = ["\ts_waitcnt vmcnt(0)", "\tglobal_load_lds_dwordx4 v[0:1], off", "\ts_barrier"]
    assert barrier_violations(late) == [2]
    # a wait that leaves vmcnt open
    lgkm = ["\tglobal_load_lds_dwordx4 v[0:1], off", "\ts_waitcnt lgkmcnt(0)", "\ts_barrier"]
    assert barrier_violations(lgkm) == [2]
    # the wait sits in another block
    other = ["\ts_waitcnt vmcnt(0)", ".LBB0_3:", "\ts_barrier"]
    assert barrier_violations(other) == [2]
    # M0 written right before the copy (a comment line between does not count)
    m0 = ["\ts_add_u32 m0, s4, 0x2000", "\t; sched_group", "\tglobal_load_lds_dwordx4 v[0:1], off"]
    assert m0_violations(m0) == [1]
