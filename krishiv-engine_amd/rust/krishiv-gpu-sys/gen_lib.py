"""Writes src/lib.rs of the krishiv-gpu-sys crate from include/kgpu.h: every
#define constant, #[repr(C)] struct, opaque handle and extern "C" function of
the header, in declaration order (what bindgen would emit for it; bindgen
is not a dependency of the build).

    python krishiv-engine_amd/rust/krishiv-gpu-sys/gen_lib.py
"""
from __future__ import annotations

import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
HEADER = os.path.join(ROOT, "include", "kgpu.h")
OUT = os.path.join(HERE, "src", "lib.rs")

SCALAR = {"int": "c_int", "int32_t": "i32", "int64_t": "i64", "uint32_t": "u32",
          "uint64_t": "u64", "uint8_t": "u8", "double": "f64", "size_t": "usize",
          "char": "c_char", "void": "c_void"}


def strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def rust_type(ctype: str) -> str:
    """A C parameter or field type (without its name) as a Rust type: each
    pointer is *const when what it points to is const."""
    toks = ctype.replace("*", " * ").split()
    if "*" not in toks:
        base = " ".join(w for w in toks if w not in ("const", "struct"))
        rb = SCALAR.get(base, base)
        return "()" if rb == "c_void" else rb
    first = toks.index("*")
    base_toks = toks[:first]
    base = " ".join(w for w in base_toks if w not in ("const", "struct"))
    t = SCALAR.get(base, base)
    pointee_const = "const" in base_toks
    rest = toks[first:]
    for k, tok in enumerate(rest):
        if tok != "*":
            continue
        t = ("*const " if pointee_const else "*mut ") + t
        pointee_const = k + 1 < len(rest) and rest[k + 1] == "const"
    return t


def field_line(f: str) -> str:
    """One struct field as a Rust field (arrays and function pointers too)."""
    fp = re.match(r"(.*?)\(\s*\*\s*(\w+)\s*\)\s*\((.*)\)$", f)
    if fp:
        ret, fname, args = fp.groups()
        params = [rust_type(a) for a in args.split(",") if a.strip() and a.strip() != "void"]
        rr = rust_type(ret)
        sig = f"unsafe extern \"C\" fn({', '.join(params)})" + ("" if rr == "()" else f" -> {rr}")
        return f"    pub {ident(fname)}: Option<{sig}>,"
    arr = re.match(r"(.*?)(\w+)\s*\[(\d+)\]$", f)
    if arr:
        ft, fname, n = arr.groups()
        return f"    pub {ident(fname)}: [{rust_type(ft)}; {n}],"
    ft, fname = split_param(f)
    return f"    pub {ident(fname)}: {rust_type(ft)},"


RUST_KEYWORDS = {"type", "in", "fn", "ref", "match", "loop", "move", "where", "use", "mod",
                 "impl", "self", "box", "crate", "let", "mut", "pub", "static", "struct",
                 "trait", "unsafe", "as", "break", "const", "continue", "else", "enum", "extern",
                 "for", "if", "return", "super", "true", "false", "while", "dyn", "async",
                 "await", "override", "final", "macro", "priv", "typeof", "yield", "do"}


def ident(name: str) -> str:
    return "r#" + name if name in RUST_KEYWORDS else name


def split_param(p: str):
    """'const kg_column *cols' -> (type, name)."""
    p = p.strip()
    m = re.match(r"(.*?)([A-Za-z_][A-Za-z0-9_]*)\s*$", p)
    return m.group(1), m.group(2)


def generate() -> str:
    src = strip_comments(open(HEADER).read())
    out = ["// Generated by gen_lib.py from include/kgpu.h: the FFI of the MI355X path",
           "// (libkgpu.so).  Do not edit by hand; regenerate after changing the header.",
           "#![allow(non_camel_case_types, non_upper_case_globals, dead_code)]",
           "",
           "use std::os::raw::{c_char, c_int, c_void};",
           ""]
    for m in re.finditer(r"#define\s+((?:KG|ARROW)_[A-Z0-9_]+)\s+([0-9]+)(u?)\b", src):
        name, val, uns = m.groups()
        out.append(f"pub const {name}: {'u32' if uns else 'i32'} = {val};")
    out.append("")
    pos = 0
    items = []
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S):
        items.append((m.start(), "struct", m))
    for m in re.finditer(r"typedef\s+struct\s+(\w+)\s+(\w+)\s*;", src):
        items.append((m.start(), "opaque", m))
    for m in re.finditer(r"(?<!typedef )\bstruct\s+(\w+)\s*\{(.*?)\}\s*;", src, flags=re.S):
        items.append((m.start(), "cstruct", m))
    for m in re.finditer(r"typedef\s+(int32_t|int64_t|uint32_t|uint64_t)\s+(\w+)\s*;", src):
        items.append((m.start(), "alias", m))
    decl = re.compile(r"^\s*([A-Za-z_][\w\s\*]*?[\s\*])(kg_\w+)\s*\(([^;]*?)\)\s*;", flags=re.M | re.S)
    for m in decl.finditer(src):
        items.append((m.start(), "fn", m))
    items.sort(key=lambda x: x[0])
    fns = []
    for _, kind, m in items:
        if kind == "opaque":
            out.append(f"#[repr(C)]\npub struct {m.group(2)} {{\n    _private: [u8; 0],\n}}\n")
        elif kind == "alias":
            out.append(f"pub type {m.group(2)} = {SCALAR[m.group(1)]};\n")
        elif kind in ("struct", "cstruct"):
            body = m.group(2)
            name = m.group(3) if kind == "struct" else m.group(1)
            fields = [field_line(" ".join(f.split())) for f in body.split(";") if f.strip()]
            out.append("#[repr(C)]\n#[derive(Clone, Copy, Debug)]\npub struct " + name +
                       " {\n" + "\n".join(fields) + "\n}\n")
        else:
            ret, name, args = m.group(1), m.group(2), " ".join(m.group(3).split())
            params = []
            if args and args != "void":
                for p in args.split(","):
                    pt, pn = split_param(p)
                    params.append(f"{ident(pn)}: {rust_type(pt)}")
            rr = rust_type(ret)
            sig = f"    pub fn {name}({', '.join(params)})" + ("" if rr == "()" else f" -> {rr}") + ";"
            fns.append(sig)
    out.append('#[link(name = "kgpu")]\nextern "C" {')
    out.extend(fns)
    out.append("}")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    text = generate()
    with open(OUT, "w") as f:
        f.write(text)
    print(f"wrote {OUT}")
