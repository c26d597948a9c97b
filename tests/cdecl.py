"""Small parsers for the boundary checks: C prototypes of include/*.h, the
`extern "C"` block and public functions of the Rust crate, and the C -> Rust
type mapping a `-sys` crate uses (size_t = usize, int = c_int, T* = *mut T,
const T* = *const T, void = c_void)."""
from __future__ import annotations

import re

_C_SCALARS = {
    "size_t": "usize", "double": "f64", "float": "f32", "bool": "bool", "int": "c_int",
    "void": "c_void", "char": "c_char", "uint8_t": "u8", "int8_t": "i8", "uint16_t": "u16",
    "int16_t": "i16", "int32_t": "i32", "uint32_t": "u32", "int64_t": "i64", "uint64_t": "u64",
}


def _strip_c(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    return text


def c_prototypes(text: str) -> dict:
    """name -> (return type, [(param type, param name)]) for every ferray_*
    function prototype."""
    text = _strip_c(text)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(ferray_\w+)\s*\(([^)]*)\)\s*;", text):
        ret = " ".join(m.group(1).split())
        if ret.startswith("typedef") or "struct" in ret:
            continue
        params = []
        body = " ".join(m.group(3).split())
        if body and body != "void":
            for p in body.split(","):
                p = p.strip()
                pm = re.match(r"(.*?)(\w+)$", p)
                params.append((" ".join(pm.group(1).replace("*", " * ").split()), pm.group(2)))
        out[m.group(2)] = (ret, params)
    return out


def c_to_rust(ctype: str) -> str:
    """A C type of the header as the -sys crate spells it."""
    toks = ctype.replace("*", " * ").split()
    stars = toks.count("*")
    base = [t for t in toks if t not in ("*",)]
    const = "const" in base
    base = [t for t in base if t != "const"]
    assert len(base) == 1, ctype
    name = _C_SCALARS.get(base[0], base[0])
    if stars == 0:
        return "()" if name == "c_void" else name
    # the first star binds to the (possibly const) base type
    rust = f"*{'const' if const else 'mut'} {name}"
    for _ in range(stars - 1):
        rust = f"*mut {rust}"
    return rust


def rust_norm(t: str) -> str:
    t = " ".join(t.replace("*", " * ").split())
    return t.replace("* ", "*")


def rust_extern_fns(text: str) -> dict:
    """name -> (return type, [(param name, type)]) inside extern "C" blocks."""
    out = {}
    for block in re.finditer(r'extern\s+"C"\s*\{(.*?)\n\}', text, flags=re.S):
        body = re.sub(r"//[^\n]*", " ", block.group(1))
        for m in re.finditer(r"pub\s+fn\s+(\w+)\s*\((.*?)\)\s*(->\s*([^;]+))?;", body, flags=re.S):
            params = []
            for p in " ".join(m.group(2).split()).split(","):
                p = p.strip()
                if not p:
                    continue
                name, ty = p.split(":", 1)
                params.append((name.strip(), rust_norm(ty)))
            ret = rust_norm(m.group(4)) if m.group(4) else "()"
            out[m.group(1)] = (ret, params)
    return out


def rust_pub_fns(text: str, module: str) -> dict:
    """name -> signature text of `pub [unsafe] fn` items in `pub mod module`."""
    m = re.search(r"pub mod " + module + r"\s*\{(.*?)\n\}", text, flags=re.S)
    assert m, module
    out = {}
    for f in re.finditer(r"(pub (?:unsafe )?fn (\w+)\s*\(.*?\)\s*(?:->\s*[^{]+)?)\{", m.group(1),
                         flags=re.S):
        out[f.group(2)] = f.group(1)
    return out


def sig_norm(sig: str) -> str:
    """Signature text with all spacing and trailing commas removed."""
    s = re.sub(r"\s+", "", sig)
    return s.replace(",)", ")")
