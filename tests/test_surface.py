"""Signature freeze of the drop-in boundary (CPU only).

  * include/ferray_gemm.h and the ferray-gemm-hip-sys crate's extern "C"
    block declare the same functions with the same parameter names and the
    C -> Rust mapping of every type;
  * the crate's safe `gemm` module keeps the signature of every public
    ferray_linalg::gemm entry it replaces, token for token, against the
    reference's own conformance snapshot (tests/golden/gemm_surface.json,
    extracted from ferray-linalg/tests/conformance/_surface.json);
    cpu_supports_* stay ferray's, and the device probe is gpu_supports;
  * each drop-in C entry takes the reference entry's parameters in order
    (slices and raw pointers as C pointers, half::bf16 / f16 as uint16_t),
    and each row-strided device entry takes the parameters of the driver it
    mirrors (outer::gemm_avx2_*) in order, followed by the stream.
"""
from __future__ import annotations

import json
import os
import re

from tests.cdecl import c_prototypes, c_to_rust, rust_extern_fns, rust_norm, rust_pub_fns, sig_norm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ferray_gemm.h")
CRATE = os.path.join(ROOT, "ferray-gemm-hip-sys", "src", "lib.rs")
SURFACE = os.path.join(ROOT, "tests", "golden", "gemm_surface.json")


def _header():
    return c_prototypes(open(HEADER).read())


def _surface():
    with open(SURFACE) as f:
        return json.load(f)


def _ref_params(sig: str):
    """[(name, type)] of a snapshot signature, types in the crate's spelling."""
    body = re.search(r"\((.*)\)", sig).group(1)
    out = []
    for p in body.split(","):
        p = p.strip()
        if not p:
            continue
        name, ty = p.split(":", 1)
        out.append((name.strip(), rust_norm(ty.replace(" :: ", "::"))))
    return out


def test_crate_extern_block_matches_the_header():
    hdr = _header()
    ext = rust_extern_fns(open(CRATE).read())
    assert set(ext) == set(hdr), (sorted(set(hdr) - set(ext)), sorted(set(ext) - set(hdr)))
    for name, (ret, params) in hdr.items():
        rret, rparams = ext[name]
        assert rret == rust_norm(c_to_rust(ret)), (name, rret, ret)
        assert [n for _, n in params] == [n for n, _ in rparams], name
        for (ct, pn), (_, rt) in zip(params, rparams):
            assert rt == rust_norm(c_to_rust(ct)), (name, pn, ct, rt)


def test_crate_keeps_the_reference_signatures():
    pub = rust_pub_fns(open(CRATE).read(), "gemm")
    for e in _surface()["public"]:
        name = e["path"].rsplit("::", 1)[1]
        if name.startswith("cpu_supports_"):
            assert name not in pub      # ferray's own CPU probes are not replaced
            continue
        assert name in pub, name
        assert sig_norm(pub[name]) == sig_norm(e["signature"]), (pub[name], e["signature"])
    assert "gpu_supports" in pub and "offload_min_dim" in pub


def _as_c_side(rust_type: str) -> str:
    """The reference parameter type as the C ABI receives it."""
    t = rust_type.replace(" ", "")
    t = t.replace("&mut[", "*mut").replace("&[", "*const").rstrip("]")
    t = t.replace("half::bf16", "u16").replace("half::f16", "u16")
    return re.sub(r"^\*(const|mut)", r"*\1 ", t)


def test_drop_in_entries_take_the_reference_parameters():
    hdr = _header()
    for e in _surface()["public"]:
        name = e["path"].rsplit("::", 1)[1]
        if name.startswith("cpu_supports_"):
            continue
        ret, params = hdr[f"ferray_{name}"]
        assert ret == "bool", name
        ref = _ref_params(e["signature"])
        assert [n for _, n in params] == [n for n, _ in ref], name
        for (ct, pn), (_, rt) in zip(params, ref):
            assert rust_norm(c_to_rust(ct)) == _as_c_side(rt), (name, pn, ct, rt)


def test_strided_entries_take_the_driver_parameters():
    hdr = _header()
    for e in _surface()["drivers"]:
        dt = e["path"].rsplit("_", 1)[1]                 # f64 / f32 / c64 / c32
        ret, params = hdr[f"ferray_gemm_{dt}_strided_dev"]
        assert ret == "int" and params[-1] == ("void *", "stream"), dt
        ref = _ref_params(e["signature"])
        assert [n for _, n in params[:-1]] == [n for n, _ in ref], dt
        for (ct, pn), (_, rt) in zip(params[:-1], ref):
            assert rust_norm(c_to_rust(ct)) == rust_norm(rt), (dt, pn, ct, rt)
