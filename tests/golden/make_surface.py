"""Writes tests/golden/gemm_surface.json: the frozen signatures of ferray's
public GEMM entry points (ferray_linalg::gemm::*, the functions the drop-in
replaces) and of the row-strided drivers the strided device entries mirror
(ferray_linalg::gemm::outer::gemm_avx2_{f64,f32,c64,c32}), copied as data
from the reference's conformance snapshot
(ferray-linalg/tests/conformance/_surface.json).  Run in a checkout that has
the reference under /root/reference; the fixture is committed."""
from __future__ import annotations

import json
import os

SRC = "/root/reference/ferray-linalg/tests/conformance/_surface.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gemm_surface.json")

PUBLIC = ["cpu_supports_avx2_fma", "cpu_supports_avx512_bw", "cpu_supports_avx512_vnni",
          "cpu_supports_avx512f", "cpu_supports_neon", "gemm_bf16_f32", "gemm_c32", "gemm_c64",
          "gemm_f16_f32", "gemm_f32", "gemm_f64", "gemm_i16", "gemm_i8", "gemm_i8_signed",
          "quantized_matmul"]
DRIVERS = ["gemm_avx2_f64", "gemm_avx2_f32", "gemm_avx2_c64", "gemm_avx2_c32"]


def main() -> None:
    with open(SRC) as f:
        items = json.load(f)["items"]
    by_path = {e["path"]: e for e in items}
    out = {"source": "ferray-linalg/tests/conformance/_surface.json",
           "public": [by_path[f"ferray_linalg::gemm::{n}"] for n in PUBLIC],
           "drivers": [by_path[f"ferray_linalg::gemm::outer::{n}"] for n in DRIVERS]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
