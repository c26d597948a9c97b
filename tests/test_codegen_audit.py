"""Audit of the compiled gfx950 code (CPU only: hipcc cross-compiles).

Kernels that copy operands with LDS-DMA loads (`global_load_lds_*`) must not
touch scratch memory, and no NT kernel may write memory with AGPR data.  On
some MI355X boards the forms that did (accumulators stored straight from
AGPRs; accumulators spilled to scratch) returned wrong words while the same
products in register-only form were exact, so the rule is checked on the
code the compiler actually emits, not on the source.
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
DMA_SOURCES = ["gemm_nt.hip", "ozaki.hip", "sgemm_split.hip"]

AGPR_WRITE = re.compile(r"^\s*(global|buffer|flat|scratch)_(store|atomic)\w*\s.*,\s*a[\[0-9]")


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


def test_audit_sees_a_planted_violation():
    """The parser flags both forms on a hand-written listing."""
    listing = "\n".join([
        "_Zfake:", "\tglobal_load_lds_dwordx4 v[0:1], off",
        "\tscratch_store_dwordx4 off, a[0:3], off", "\tglobal_store_dword v[2:3], a7, off",
        "\tglobal_store_dword v[2:3], v7, off", "\ts_endpgm"])
    body = _kernels(listing)["_Zfake"]
    assert sum("scratch_" in l for l in body) == 1
    assert sum(bool(AGPR_WRITE.match(l)) for l in body) == 2
