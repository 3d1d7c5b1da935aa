#!/usr/bin/env python3
"""Writes src/lib.rs, the raw Rust declarations of the C ABI, from include/vt3d_dense.h.

The header is the one source of truth: every #define constant, enum value, struct (as a
#[repr(C)] mirror with the same field order and types), opaque handle, the poll callback and
every function prototype is translated mechanically, so the Rust side cannot drift from the
library. Run it after changing the header:

    python video-to-3d_amd/rust/vt3d-dense-sys/generate.py            # rewrite src/lib.rs
    python video-to-3d_amd/rust/vt3d-dense-sys/generate.py --check    # exit 1 if stale
"""
from __future__ import annotations

import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.normpath(os.path.join(HERE, "..", "..", "..", "include", "vt3d_dense.h"))
OUT = os.path.join(HERE, "src", "lib.rs")

# C scalar -> Rust (x86_64 / aarch64 LP64: size_t is usize, char is c_char).
SCALARS = {
    "uint8_t": "u8", "int8_t": "i8", "uint16_t": "u16", "int16_t": "i16",
    "uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "int64_t": "i64",
    "float": "f32", "double": "f64", "size_t": "usize", "char": "c_char", "void": "c_void",
}


def strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def rust_type(ctype: str, stars: int, structs: set) -> str:
    """`const T` / `T` with `stars` pointer levels -> the Rust type."""
    const = ctype.startswith("const ")
    base = ctype[len("const "):] if const else ctype
    base = base.strip()
    if base in SCALARS:
        r = SCALARS[base]
    elif base in structs or base == "vt3d_poll_fn":
        r = base
    else:
        raise ValueError(f"unknown C type {ctype!r}")
    for level in range(stars):
        # the innermost pointer carries the const of the pointee
        r = f"*{'const' if const and level == 0 else 'mut'} {r}"
    return r


def split_decl(decl: str):
    """'const uint8_t *rgba' -> ('const uint8_t', 1, 'rgba', None); arrays keep their length
    expression: 'double t[3]' -> ('double', 0, 't', '3')."""
    decl = " ".join(decl.split())
    m = re.match(r"^(.*?)([\s\*]+)(\w+)\s*(?:\[([^\]]+)\])?$", decl)
    if not m:
        raise ValueError(f"cannot parse declaration {decl!r}")
    base, sep, name, dim = m.groups()
    return base.strip(), sep.count("*"), name, dim


def parse(text: str):
    raw = text
    text = strip_comments(text)
    defines = re.findall(r"^#define\s+(VT3D_\w+)\s+(\(?-?\d+\)?)\s*$", text, flags=re.M)
    defines = [(n, v.strip("()")) for n, v in defines]
    enums = []
    for name, body in re.findall(r"enum\s+(\w+)\s*\{(.*?)\}\s*;", text, flags=re.S):
        for item in body.split(","):
            item = item.strip()
            if item:
                k, v = (x.strip() for x in item.split("="))
                enums.append((name, k, v))
    opaque = re.findall(r"typedef\s+struct\s+(\w+)\s+\1\s*;", text)
    structs = []
    for name, body in re.findall(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*\1\s*;", text, flags=re.S):
        fields = []
        for stmt in body.split(";"):
            stmt = " ".join(stmt.split())
            if not stmt:
                continue
            # 'float x, y, z, confidence' shares one base type
            first, *more = [p.strip() for p in stmt.split(",")]
            base, stars, fname, dim = split_decl(first)
            fields.append((fname, base, stars, dim))
            for extra in more:
                _, stars2, fname2, dim2 = split_decl(base + " " + extra)
                fields.append((fname2, base, stars2, dim2))
        structs.append((name, fields))
    # functions: everything outside the typedefs that looks like `ret name(args);`
    body = re.sub(r"typedef\s+struct\s+\w+\s*\{.*?\}\s*\w+\s*;", " ", text, flags=re.S)
    body = re.sub(r"typedef[^;]*;", " ", body)
    body = re.sub(r"enum\s+\w+\s*\{.*?\}\s*;", " ", body, flags=re.S)
    body = re.sub(r"^#.*$", " ", body, flags=re.M)
    body = body.replace('extern "C" {', " ").replace("}", " ")
    funcs = []
    for ret, name, args in re.findall(r"([A-Za-z_][\w\s\*]*?)\b(vt3d_\w+)\s*\(([^)]*)\)\s*;", body):
        ret = " ".join(ret.split())
        stars = ret.count("*")
        ret = ret.replace("*", "").strip()
        params = []
        args = " ".join(args.split())
        if args and args != "void":
            for a in args.split(","):
                base, pstars, pname, dim = split_decl(a.strip())
                if dim is not None:  # an array parameter is a pointer
                    pstars += 1
                params.append((pname, base, pstars))
        funcs.append((name, ret, stars, params))
    poll = re.search(r"typedef\s+int32_t\s+\(\*vt3d_poll_fn\)\(void \*user\);", text)
    if not poll:
        raise ValueError("vt3d_poll_fn typedef changed: update the generator")
    comments = {}
    for m in re.finditer(r"/\*((?:(?!\*/).)*?)\*/\s*typedef\s+struct\s+(\w+)\s*\{", raw, flags=re.S):
        lines = [re.sub(r"^\s*\*?\s?", "", ln) for ln in m.group(1).splitlines()]
        comments[m.group(2)] = " ".join(" ".join(lines).split())
    return defines, enums, opaque, structs, funcs, comments


def render(header_text: str) -> str:
    defines, enums, opaque, structs, funcs, comments = parse(header_text)
    known = set(opaque) | {n for n, _ in structs}
    # constants that size arrays (struct fields, or array parameters) are usize
    dims = set()
    for _, fields in structs:
        for _, _, _, dim in fields:
            if dim:
                dims.update(re.findall(r"VT3D_\w+", dim))
    dims.update(re.findall(r"\[\s*(VT3D_\w+)\s*\]", strip_comments(header_text)))
    out = [
        "// @generated by generate.py from include/vt3d_dense.h -- do not edit by hand.",
        "//",
        "// Raw declarations of libvt3d_dense.so, the MI355X dense plane-sweep depth pass. The",
        "// safe replacement of dense_legacy::estimate_depth_points_with_control",
        "// (crates/video-to-3d-core/src/dense_legacy.rs:149-156) and of the other entry points",
        "// is written on top of these in video-to-3d-core.",
        "#![no_std]",
        "#![allow(non_camel_case_types, non_upper_case_globals, non_snake_case)]",
        "",
        "pub use core::ffi::{c_char, c_void};",
        "",
    ]
    for n, v in defines:
        ty = "usize" if n in dims else "i32"
        out.append(f"pub const {n}: {ty} = {v};")
    out.append("")
    for _, k, v in enums:
        out.append(f"pub const {k}: i32 = {v};")
    out.append("")
    out.append("/// ReconstructionControl poll: nonzero cancels (called on the calling thread).")
    out.append("pub type vt3d_poll_fn = Option<unsafe extern \"C\" fn(user: *mut c_void) -> i32>;")
    out.append("")
    for name in opaque:
        out.append(f"/// Opaque library handle.")
        out.append("#[repr(C)]")
        out.append(f"pub struct {name} {{")
        out.append("    _private: [u8; 0],")
        out.append("}")
        out.append("")
    for name, fields in structs:
        if name in comments:
            out.append(f"/// {comments[name]}")
        out.append("#[repr(C)]")
        out.append("#[derive(Clone, Copy)]")
        out.append(f"pub struct {name} {{")
        for fname, base, stars, dim in fields:
            ty = rust_type(base, stars, known)
            if dim is not None:
                n = dim.replace(" ", "")
                ty = f"[{ty}; {n}]"
            out.append(f"    pub {fname}: {ty},")
        out.append("}")
        out.append("")
    out.append("extern \"C\" {")
    for name, ret, rstars, params in funcs:
        ps = ", ".join(f"{p}: {rust_type(b, s, known)}" for p, b, s in params)
        r = "" if ret == "void" and rstars == 0 else f" -> {rust_type(ret, rstars, known)}"
        out.append(f"    pub fn {name}({ps}){r};")
    out.append("}")
    return "\n".join(out) + "\n"


def main() -> int:
    text = render(open(HEADER).read())
    if "--check" in sys.argv[1:]:
        current = open(OUT).read() if os.path.exists(OUT) else ""
        if current != text:
            print(f"{OUT} is stale: run {os.path.relpath(__file__)}", file=sys.stderr)
            return 1
        return 0
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
