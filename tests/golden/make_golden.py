"""Writes the golden fixtures under tests/golden/ (run in the build container).

* matmul.json / dot.json: the reference's own numpy-generated fixtures
  (fixtures/linalg/{matmul,dot}.json, numpy 2.4.5), copied as data.
* ferray_kats.json: the known-answer vectors held in the reference's tests,
  transcribed as data:
    - gemm_f64 2x2 -> [19, 22, 43, 50]  (ferray-linalg/tests/conformance_internal_surface.rs:741-746)
    - numpy hex goldens: cancellation 2x3x2 exact, random 4x3x5 <= 2 ulp
      (ferray-linalg/src/gemm/mod.rs:1213-1275)
    - matmul_complex 2x2 identity and non-trivial cases (ferray-linalg/src/complex.rs:729-782)
The reference is read here only as data; nothing in it is executed.
"""
import json
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def main() -> None:
    for name in ("matmul.json", "dot.json"):
        shutil.copyfile(os.path.join(REF, "fixtures", "linalg", name), os.path.join(HERE, name))
    kats = {
        "gemm_f64_2x2": {"m": 2, "n": 2, "k": 2, "a": [1.0, 2.0, 3.0, 4.0],
                         "b": [5.0, 6.0, 7.0, 8.0], "c": [19.0, 22.0, 43.0, 50.0],
                         "source": "ferray-linalg/tests/conformance_internal_surface.rs:741-746"},
        "numpy_cancellation_2x3x2": {
            "m": 2, "n": 2, "k": 3, "max_ulp": 0,
            "a": ["0x1.7d78400000000p+26", "0x1.0000000000000p+0", "-0x1.7d78400000000p+26",
                  "0x1.8000000000000p+1", "-0x1.0000000000000p+1", "0x1.8000000000000p+0"],
            "b": ["0x1.0p+0", "0x1.0p+1", "0x1.0p+0", "0x1.0p+0", "0x1.0p+0", "0x1.0p+1"],
            "c": ["0x1.0000000000000p+0", "0x1.0000000000000p+0", "0x1.4000000000000p+1",
                  "0x1.c000000000000p+2"],
            "source": "ferray-linalg/src/gemm/mod.rs:1213-1242"},
        "numpy_random_4x3x5": {
            "m": 4, "n": 5, "k": 3, "max_ulp": 2,
            "a": ["0x1.427a31c1ffc58p-10", "0x1.31ea59a5b303ep-2", "-0x1.18b7980d81558p-2",
                  "-0x1.c7fba74b18102p-1", "-0x1.d19537e309692p-2", "-0x1.fbb918e5cd3f0p-1",
                  "0x1.ecb246c7071e5p-5", "0x1.571858a941e07p+0", "-0x1.f804fc5039525p-2",
                  "-0x1.3daee2d56e58bp-1", "0x1.f5992787010aap-2", "0x1.6d73c9b1a8b33p-2"],
            "b": ["0x1.afc6d9ffa76cap-4", "-0x1.dc664ebbfd571p-1", "-0x1.df43093500899p-6",
                  "0x1.63fec7c2015e9p-1", "-0x1.581e71cf65997p+0",
                  "-0x1.d49939df35222p-2", "-0x1.e6b68891db2c3p+0", "-0x1.4a1f253355902p+0",
                  "-0x1.d77bf28b367b0p+0", "-0x1.e177757c5cfd8p-3",
                  "-0x1.44775f633b214p+0", "0x1.15c652f6d6a5ep-2", "0x1.4106b6b54ede6p-3",
                  "-0x1.7ed5a6ae5e4abp-3", "-0x1.42252ea4eea8ap+1"],
            "c": ["0x1.afdebe7d16fb3p-3", "-0x1.49778e149b576p-1", "-0x1.b687708af82bfp-2",
                  "-0x1.fe10889f664f1p-2", "0x1.3c71747d3dda5p-1",
                  "0x1.5efca23ada43dp+0", "0x1.6c91c5d3ada6ep+0", "0x1.d3e41eb3a5f83p-2",
                  "0x1.9d34a8d49f5fcp-2", "0x1.e65eeb0aa9f91p+1",
                  "0x1.14975f74502e8p-6", "-0x1.5e674cfa638d3p+1", "-0x1.cea2c82ccec68p+0",
                  "-0x1.2ad0b8fa93377p+1", "0x1.af89a16362a42p-1",
                  "-0x1.7bda84228ddf6p-1", "-0x1.0753ea8d135ffp-2", "-0x1.1d7ac7b02b82dp-1",
                  "-0x1.66797312692e5p+0", "-0x1.6f3772d5dd5b1p-3"],
            "source": "ferray-linalg/src/gemm/mod.rs:1244-1275"},
        "complex_2x2_identity": {
            "m": 2, "n": 2, "k": 2, "tol": 0.0,
            "a": [[1, 1], [2, 0], [3, 0], [4, -1]], "b": [[1, 0], [0, 0], [0, 0], [1, 0]],
            "c": [[1, 1], [2, 0], [3, 0], [4, -1]],
            "source": "ferray-linalg/src/complex.rs:729-749"},
        "complex_2x2_nontrivial": {
            "m": 2, "n": 2, "k": 2, "tol": 1e-12,
            "a": [[1, 1], [2, -1], [3, 0], [1, 0]], "b": [[1, 0], [0, -1], [0, 1], [1, 0]],
            "c": [[2, 3], [3, -2], [3, 1], [1, -3]],
            "source": "ferray-linalg/src/complex.rs:752-782"},
    }
    with open(os.path.join(HERE, "ferray_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)


if __name__ == "__main__":
    main()
