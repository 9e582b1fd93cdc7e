"""The Rust crate's declarations (visloc-rs_amd/rust/src/ffi.rs) against the
C headers it binds (include/vba.h, include/vba_io.h). There is no Rust
toolchain here, so the layout `#[repr(C)]` gives each struct is computed from
the declared field types by C's rules and compared with what gcc reports for
the header's struct: field names in order, every offset, every size. Every
constant must equal the header's value, every function of the headers must be
declared with the same parameter count, and the struct literals the crate's
sources build must name exactly the fields of the struct."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "visloc-rs_amd", "rust")
HEADERS = [os.path.join(ROOT, "include", "vba.h"), os.path.join(ROOT, "include", "vba_io.h")]
PRIM = {"u8": 1, "i8": 1, "u16": 2, "i16": 2, "u32": 4, "i32": 4, "f32": 4, "u64": 8, "i64": 8, "f64": 8,
        "usize": 8, "isize": 8}


def read(path):
    with open(path) as f:
        return f.read()


def strip_rust_comments(src):
    return re.sub(r"//[^\n]*", "", src)


def rust_structs(src):
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+)\s*\{(.*?)\n\}", src, re.S):
        fields = re.findall(r"pub (\w+): ([^,\n]+),", m.group(2))
        if fields:
            out[m.group(1)] = [(n, t.strip()) for n, t in fields]
    return out


def rust_consts(src):
    return {n: int(eval(v.replace("u32", "").replace("i32", ""))) for n, _, v in
            re.findall(r"pub const (\w+): (\w+) = ([^;]+);", src)}


def rust_functions(src):
    out = {}
    for block in re.findall(r'extern "C" \{(.*?)\n\}', src, re.S):
        for name, params in re.findall(r"pub fn (\w+)\((.*?)\)", block, re.S):
            out[name] = len(re.findall(r"\b\w+\s*:", params))
    return out


def c_functions(src):
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(vba_\w+)\s*\(([^()]*)\)\s*;", src):
        p = params.strip()
        out[name] = 0 if p in ("", "void") else p.count(",") + 1
    return out


def layout(structs, name, memo):
    """(size, align, [(field, offset, size)]) of a repr(C) struct."""
    if name in memo:
        return memo[name]
    off, align, fields = 0, 1, []
    for fname, ftype in structs[name]:
        size, a = type_layout(structs, ftype, memo)
        off = (off + a - 1) // a * a
        fields.append((fname, off, size))
        off += size
        align = max(align, a)
    size = (off + align - 1) // align * align
    memo[name] = (size, align, fields)
    return memo[name]


def type_layout(structs, t, memo):
    t = t.strip()
    if t.startswith("*const") or t.startswith("*mut"):
        return 8, 8
    m = re.fullmatch(r"\[(.+);\s*(\d+)\]", t)
    if m:
        size, align = type_layout(structs, m.group(1), memo)
        return size * int(m.group(2)), align
    if t in PRIM:
        return PRIM[t], PRIM[t]
    size, align, _ = layout(structs, t, memo)
    return size, align


@pytest.fixture(scope="module")
def ffi():
    return strip_rust_comments(read(os.path.join(CRATE, "src", "ffi.rs")))


@pytest.fixture(scope="module")
def c_layout(ffi, tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is needed to read the C layout")
    structs = rust_structs(ffi)
    consts = rust_consts(ffi)
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "vba.h"', '#include "vba_io.h"',
             "int main(void) {"]
    for s, fields in structs.items():
        lines.append(f'    printf("size {s} %zu\\n", sizeof({s}));')
        for f, _ in fields:
            lines.append(f'    printf("field {s} {f} %zu %zu\\n", offsetof({s}, {f}), sizeof((({s}*)0)->{f}));')
    for c in consts:
        lines.append(f'    printf("const {c} %lld\\n", (long long)({c}));')
    lines += ["    return 0;", "}"]
    d = tmp_path_factory.mktemp("ffi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes, fields, values = {}, {}, {}
    for line in out:
        p = line.split()
        if not p:
            continue
        if p[0] == "size":
            sizes[p[1]] = int(p[2])
        elif p[0] == "field":
            fields[(p[1], p[2])] = (int(p[3]), int(p[4]))
        elif p[0] == "const":
            values[p[1]] = int(p[2])
    return sizes, fields, values


def test_every_struct_has_the_header_layout(ffi, c_layout):
    sizes, fields, _ = c_layout
    structs = rust_structs(ffi)
    header = "\n".join(read(h) for h in HEADERS)
    c_struct_names = set(re.findall(r"typedef struct (vba_\w+) \{", header))
    assert c_struct_names == set(structs), (c_struct_names ^ set(structs))
    memo = {}
    for s in structs:
        size, _, rust_fields = layout(structs, s, memo)
        body = re.search(r"typedef struct " + s + r" \{(.*?)\} " + s + ";", header, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        c_names = re.findall(r"(\w+)(?:\[\d+\])?;", body)
        assert [f for f, _ in structs[s]] == c_names, s
        assert size == sizes[s], (s, size, sizes[s])
        for f, off, fsize in rust_fields:
            assert (off, fsize) == fields[(s, f)], (s, f, off, fsize, fields[(s, f)])


def test_every_constant_has_the_header_value(ffi, c_layout):
    _, _, values = c_layout
    consts = rust_consts(ffi)
    assert consts and all(consts[c] == values[c] for c in consts), \
        {c: (consts[c], values[c]) for c in consts if consts[c] != values[c]}
    header = "\n".join(read(h) for h in HEADERS)
    declared = set(re.findall(r"^\s*(VBA_[A-Z0-9_]+)\s*=", header, re.M)) | \
        set(re.findall(r"#define (VBA_[A-Z0-9_]+) ", header))
    declared.discard("VBA_H")
    assert declared <= set(consts), declared - set(consts)


def test_every_function_is_declared_with_its_parameter_count(ffi):
    c = {}
    for h in HEADERS:
        c.update(c_functions(read(h)))
    rust = rust_functions(ffi)
    assert set(c) == set(rust), set(c) ^ set(rust)
    assert {n: rust[n] for n in c if rust[n] != c[n]} == {}


def test_the_library_exports_every_function_the_crate_binds(ffi):
    so = os.path.join(ROOT, "visloc-rs_amd", "libvba.so")
    if not os.path.exists(so) or shutil.which("nm") is None:
        pytest.skip("libvba.so is not built")
    exported = set(subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                                  check=True).stdout.split())
    missing = [f for f in rust_functions(ffi) if f not in exported]
    assert missing == []


def balanced(src, start):
    depth = 0
    for i in range(start, len(src)):
        if src[i] == "{":
            depth += 1
        elif src[i] == "}":
            depth -= 1
            if depth == 0:
                return src[start + 1:i]
    raise AssertionError("unbalanced braces")


def top_level_keys(body):
    """Field names of a struct literal body: `name: expr` and the shorthand
    `name`, split at top-level commas."""
    items, depth, cur = [], 0, ""
    for ch in body:
        if ch in "{([":
            depth += 1
        elif ch in "})]":
            depth -= 1
        if ch == "," and depth == 0:
            items.append(cur)
            cur = ""
        else:
            cur += ch
    items.append(cur)
    keys = []
    for item in (i.strip() for i in items):
        m = re.match(r"(\w+)\s*:(?!:)", item)
        if m:
            keys.append(m.group(1))
        elif re.fullmatch(r"\w+", item):
            keys.append(item)
    return keys


def test_struct_literals_in_the_crate_name_every_field(ffi):
    structs = rust_structs(ffi)
    checked = 0
    for fn in sorted(os.listdir(os.path.join(CRATE, "src"))):
        src = strip_rust_comments(read(os.path.join(CRATE, "src", fn)))
        for m in re.finditer(r"(?<![\w:])ffi::(vba_\w+)\s*\{", src):
            if re.search(r"->\s*&?\s*$", src[:m.start()]):  # a return type, not a literal
                continue
            name = m.group(1)
            body = balanced(src, m.end() - 1)
            keys = top_level_keys(body)
            if ".." in body.replace("..=", ""):  # struct update syntax: a subset
                assert set(keys) <= {f for f, _ in structs[name]}, (fn, name)
            else:
                assert sorted(keys) == sorted(f for f, _ in structs[name]), (fn, name, keys)
            checked += 1
    assert checked >= 6


def test_the_crate_refers_only_to_declared_items(ffi):
    declared = set(rust_structs(ffi)) | set(rust_consts(ffi)) | set(rust_functions(ffi)) | \
        set(re.findall(r"pub struct (\w+)", ffi))
    for fn in sorted(os.listdir(os.path.join(CRATE, "src"))):
        if fn == "ffi.rs":
            continue
        used = set(re.findall(r"(?<![\w:])ffi::(\w+)", strip_rust_comments(read(os.path.join(CRATE, "src", fn)))))
        assert used <= declared, (fn, used - declared)
