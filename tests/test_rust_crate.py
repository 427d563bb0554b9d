"""The flashraw-hip crate (crates/flashraw-hip) against the C ABI, without a Rust toolchain: every
#[repr(C)] struct of src/ffi.rs has the header's fields in the header's order with the offsets and
sizes gcc gives them (Rust's repr(C) layout rules, computed here), every extern function matches a
header prototype (name and argument count) and is exported by libflashraw_hip.so, and the status,
input, output, format and precision constants equal the header's enumerators."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FFI = os.path.join(ROOT, "crates", "flashraw-hip", "src", "ffi.rs")
HDR = os.path.join(ROOT, "include", "flashraw_hip.h")
LIB = os.path.join(ROOT, "flashraw_amd", "lib", "libflashraw_hip.so")

PRIM = {"u8": 1, "i8": 1, "u16": 2, "i16": 2, "u32": 4, "i32": 4, "f32": 4, "u64": 8, "i64": 8, "f64": 8,
        "usize": 8, "isize": 8}


def rust_structs():
    src = open(FFI).read()
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\][^\n]*\n(?:#\[[^\n]*\n)*pub struct (\w+) \{(.*?)\n\}", src, re.S):
        fields = re.findall(r"pub (\w+): ([^,]+),", body)
        if fields:
            out[name] = [(f, t.strip()) for f, t in fields]
    return out


def layout(ty, structs, cache):
    """(size, align) of a Rust type under repr(C) on x86-64."""
    if ty in PRIM:
        return PRIM[ty], PRIM[ty]
    if ty.startswith("*") or ty.startswith("Option<"):
        return 8, 8
    m = re.fullmatch(r"\[(\w+); (\d+)\]", ty)
    if m:
        s, a = layout(m.group(1), structs, cache)
        return s * int(m.group(2)), a
    if ty in structs:
        return struct_layout(ty, structs, cache)[0]
    raise AssertionError(f"unknown type {ty}")


def struct_layout(name, structs, cache):
    if name in cache:
        return cache[name]
    off, align, fields = 0, 1, []
    for f, t in structs[name]:
        s, a = layout(t, structs, cache)
        off = (off + a - 1) // a * a
        fields.append((f, off, s))
        off += s
        align = max(align, a)
    size = (off + align - 1) // align * align
    cache[name] = ((size, align), fields)
    return cache[name]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    exe = tmp_path_factory.mktemp("abi") / "abi_layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "abi_layout.c"),
                    "-o", str(exe)], check=True)
    rows = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    out = {}
    for r in rows:
        if not r.strip():
            continue
        st, field, off, size = r.split()
        out.setdefault(st, {"fields": []})
        if field == "-":
            out[st]["size"] = int(size)
        else:
            out[st]["fields"].append((field, int(off), int(size)))
    return out


def test_every_header_struct_is_in_the_crate(c_layout):
    assert set(c_layout) <= set(rust_structs())


@pytest.mark.parametrize("name", sorted(["frh_rect", "frh_frame", "frh_frame_info", "frh_edge_spec", "frh_curve",
                                         "frh_tone", "frh_develop_spec", "frh_timings", "frh_dn_spec",
                                         "frh_nr_plane", "frh_nr_spec", "frh_server_config", "frh_job_result",
                                         "frh_server_stats"]))
def test_repr_c_layout_matches_gcc(c_layout, name):
    structs = rust_structs()
    (size, _), fields = struct_layout(name, structs, {})
    assert size == c_layout[name]["size"], name
    assert fields == c_layout[name]["fields"], name


def header_prototypes():
    s = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(frh_[a-z0-9_]+)\(([^;]*?)\);", s):
        args = " ".join(args.split())
        if name == "frh_cancel_token_with_poll":
            n = 2
        else:
            n = 0 if args == "void" else args.count(",") + 1
        out[name] = n
    return out


def rust_functions():
    """name -> argument count of each `pub fn` in ffi.rs's extern block (commas at nesting depth 0
    of the argument list, so function-pointer arguments count once)."""
    s = open(FFI).read()
    block = s[s.index('unsafe extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (frh_\w+)\(", block):
        i, depth, commas, nonblank = m.end(), 1, 0, False
        while depth:
            c = block[i]
            if c in "(<":
                depth += 1
            elif c == ")" or (c == ">" and block[i - 1] != "-"):
                depth -= 1
            elif c == "," and depth == 1:
                commas += 1
            if depth and not c.isspace():
                nonblank = True
            i += 1
        out[m.group(1)] = commas + 1 if nonblank else 0
    return out


def test_functions_match_the_header():
    h, r = header_prototypes(), rust_functions()
    assert set(h) == set(r)
    for name in h:
        assert h[name] == r[name], name


@pytest.mark.skipif(not os.path.exists(LIB), reason="libflashraw_hip.so not built")
def test_functions_are_exported():
    syms = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in syms.splitlines() if line.strip()}
    missing = sorted(set(rust_functions()) - exported)
    assert not missing, missing


def test_constants_match_the_header():
    hdr = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    enums = dict((k, int(v)) for k, v in re.findall(r"\b(FRH_[A-Z0-9_]+)\s*=\s*(-?\d+)", hdr))
    consts = dict((k, int(v)) for k, v in re.findall(r"pub const (FRH_[A-Z0-9_]+): i32 = (-?\d+);", open(FFI).read()))
    assert consts and set(consts) <= set(enums)
    for k, v in consts.items():
        assert enums[k] == v, k
    # every status, input kind, output, pixel format and precision of the header has its constant
    for k in enums:
        if k.startswith(("FRH_OK", "FRH_ERROR_", "FRH_INPUT_", "FRH_OUTPUT_", "FRH_FORMAT_", "FRH_DN_", "FRH_PACK_", "FRH_KERNEL_")):
            assert k in consts, k
