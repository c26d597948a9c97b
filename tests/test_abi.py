"""CPU-side checks of the boundary: the C-ABI library loads, exports every
function include/*.h declares, the Python mirror validates buffers before
touching the device, and the product never reaches the oracle."""
from __future__ import annotations

import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared() -> set[str]:
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ferray_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_reference_surface():
    names = _declared()
    for fn in ("gemm_f64", "gemm_f32", "gemm_c64", "gemm_c32", "gemm_i8", "gemm_i8_signed",
               "gemm_i16", "quantized_matmul", "gemm_bf16_f32", "gemm_f16_f32"):
        assert f"ferray_{fn}" in names and f"ferray_{fn}_dev" in names


def test_library_exports_every_declared_symbol():
    import ferray_amd
    lib = ctypes.CDLL(ferray_amd.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ferray_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert _declared() <= exported


def test_python_binding_covers_the_header():
    from ferray_amd._lib import SIGNATURES
    assert _declared() == set(SIGNATURES)


def test_library_is_gfx950_only():
    import tempfile

    import ferray_amd
    with tempfile.TemporaryDirectory() as d:
        fat = os.path.join(d, "lib.fatbin")
        subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin",
                        ferray_amd.LIB_PATH, fat], check=True)
        out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o",
                              f"--input={fat}"], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("clang-offload-bundler could not list the bundle")
    targets = [t for t in out.stdout.split() if "amdgcn" in t]
    assert targets and all("gfx950" in t for t in targets), targets


def test_buffer_validation_happens_before_the_device():
    import ferray_amd as fa
    a = np.zeros(6); b = np.zeros(6); c = np.zeros(3)
    with pytest.raises(ValueError):
        fa.gemm_f64(2, 2, 3, 1.0, a, b, 0.0, c)        # c too short
    with pytest.raises(TypeError):
        fa.gemm_f64(2, 2, 3, 1.0, a.astype(np.float32), b, 0.0, np.zeros(4))
    with pytest.raises(ValueError):
        fa.gemm_i8(2, 2, 3, np.zeros(5, np.uint8), np.zeros(6, np.int8), np.zeros(4, np.int32),
                   False)


def test_no_device_reports_false_not_a_crash():
    """With no GPU visible the drop-in entry points return false (the
    caller's fallback signal, products/mod.rs:509-520) and say why."""
    code = ("import numpy as np, ferray_amd as fa\n"
            "c=np.zeros(4)\n"
            "ok=fa.gemm_f64(2,2,2,1.0,np.ones(4),np.ones(4),0.0,c)\n"
            "print(ok, fa.last_status(), fa.gpu_supports())\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, cwd=ROOT,
                         env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    ok, status, avail = out.stdout.split()
    assert ok == "False" and avail == "False" and int(status) != 0


def test_product_never_imports_the_oracle():
    for path in glob.glob(os.path.join(ROOT, "ferray_amd", "**", "*"), recursive=True):
        if os.path.isfile(path) and path.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
            text = open(path, errors="ignore").read()
            assert "import oracle" not in text and "from oracle" not in text, path
            assert "libferray_oracle" not in text and "ferray_oracle.c" not in text, path
