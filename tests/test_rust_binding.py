"""The Rust binding (oxideav-vp8_amd/rust) agrees with include/vp8g.h.

No Rust toolchain is needed: the test parses both sides.
* Every function the header declares appears in vp8g-sys's extern block.
  Names, order, arity, each argument's type and the return type all match,
  with C arrays compared as the pointers they decay to.
* Every VP8G_* #define is a `pub const` of the same value.
* Every struct has the same field names and types.  Offsets and sizes match
  the C compiler's (gcc offsetof/sizeof of the header) against a repr(C)
  layout computed from the Rust field types.
* build.rs compiles the same translation units with the same hipcc flags as
  oxideav-vp8_amd/build.py.
* Every vp8g_* item the safe crate calls exists in vp8g-sys.
* When libvp8g.so is built, it exports every extern name."""
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "vp8g.h"
RUST = ROOT / "oxideav-vp8_amd" / "rust"
SYS = RUST / "vp8g-sys" / "src" / "lib.rs"
SAFE = RUST / "vp8g" / "src" / "lib.rs"

C_OF_RUST = {"u8": "uint8_t", "i8": "int8_t", "u16": "uint16_t", "i16": "int16_t", "u32": "uint32_t",
             "i32": "int32_t", "u64": "uint64_t", "i64": "int64_t", "usize": "size_t", "f32": "float",
             "f64": "double", "c_int": "int", "c_char": "char", "c_void": "void"}
PRIM = {"uint8_t": 1, "int8_t": 1, "char": 1, "uint16_t": 2, "int16_t": 2, "uint32_t": 4, "int32_t": 4, "int": 4,
        "float": 4, "uint64_t": 8, "int64_t": 8, "size_t": 8, "double": 8}


def strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", "", s, flags=re.S)


def strip_rust_comments(s):
    return re.sub(r"//[^\n]*", "", s)


# ---------------------------------------------------------------- C side

def c_type(decl):
    """(base, dims, [pointer levels], const pointee) of one C declarator
    such as 'const int32_t filters[8][6]' -> name and normalized type.
    Array parameters decay: the first dimension becomes a pointer."""
    decl = decl.strip()
    m = re.match(r"^(.*?)([A-Za-z_]\w*)\s*((?:\[\s*\d+\s*\])*)$", decl)
    assert m, decl
    t, name, arr = m.group(1), m.group(2), m.group(3)
    dims = [int(d) for d in re.findall(r"\d+", arr)]
    const = "const" in t.split("*")[0].split()
    nptr = t.count("*")
    base = t.replace("*", " ").replace("const", " ").split()
    base = " ".join(base)
    return name, base, dims, nptr, const


def c_param_type(decl):
    name, base, dims, nptr, const = c_type(decl)
    if dims:  # T x[A][B] -> pointer to T[B]
        dims, nptr = dims[1:], nptr + 1
    return name, (base, tuple(dims), nptr, const if nptr else False)


def header_functions():
    src = strip_c_comments(HEADER.read_text())
    out = []
    for ret, name, args in re.findall(r"^\s*([A-Za-z_][\w\s\*]*?)\b(vp8g_\w+)\s*\(([^;]*?)\)\s*;", src, flags=re.M):
        ret = ret.strip()
        nptr = ret.count("*")
        rbase = " ".join(ret.replace("*", " ").replace("const", " ").split())
        rconst = ret.startswith("const") and nptr > 0
        params = []
        args = " ".join(args.split())
        if args and args != "void":
            params = [c_param_type(a) for a in args.split(",")]
        out.append((name, (rbase, (), nptr, rconst), params))
    return out


def header_structs():
    src = strip_c_comments(HEADER.read_text())
    out = {}
    for body, sname in re.findall(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S):
        fields = []
        for stmt in body.split(";"):
            stmt = " ".join(stmt.split())
            if not stmt:
                continue
            m = re.match(r"^((?:const\s+)?[A-Za-z_]\w*)\s+(.*)$", stmt)
            base_t, decls = m.group(1), m.group(2)
            for d in decls.split(","):
                name, base, dims, nptr, const = c_type(base_t + " " + d.strip())
                fields.append((name, (base, tuple(dims), nptr, const if nptr else False)))
        out[sname] = fields
    return out


def header_defines():
    src = strip_c_comments(HEADER.read_text())
    return {k: int(v) for k, v in re.findall(r"^#define\s+(VP8G_\w+)\s+(-?\d+)\s*$", src, flags=re.M)}


def header_opaque():
    return set(re.findall(r"typedef\s+struct\s+(\w+)\s+\1\s*;", strip_c_comments(HEADER.read_text())))


# ---------------------------------------------------------------- Rust side

def rust_type(t):
    """Normalized (base, dims, nptr, const pointee) of a Rust type; the
    const flag is the innermost pointer's, as C's `const T *...`."""
    chain = re.findall(r"\*(const|mut)\s*", t)
    t = re.sub(r"^(\*(const|mut)\s+)+", "", t.strip()).strip()
    dims = []
    while t.startswith("["):
        m = re.match(r"^\[(.*);\s*(\d+)\]$", t)
        t = m.group(1).strip()
        dims.append(int(m.group(2)))
    return (C_OF_RUST.get(t, t), tuple(dims), len(chain), bool(chain) and chain[-1] == "const")


def rust_pointer_chain(t):
    return re.findall(r"\*(const|mut)", t)


def extern_block():
    src = strip_rust_comments(SYS.read_text())
    m = re.search(r'extern\s+"C"\s*\{(.*)\}\s*$', src, flags=re.S)
    assert m, "no extern \"C\" block in vp8g-sys"
    return m.group(1)


def split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "[(":
            depth += 1
        elif ch in "])":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [a.strip() for a in out if a.strip()]


def rust_functions():
    body = extern_block()
    out = []
    for name, args, ret in re.findall(r"pub\s+fn\s+(vp8g_\w+)\s*\((.*?)\)\s*(?:->\s*([^;]+))?;", body, flags=re.S):
        params = []
        for a in split_args(" ".join(args.split())):
            n, t = a.split(":", 1)
            n = n.strip().removeprefix("r#")
            params.append((n, rust_type(t), rust_pointer_chain(t)))
        r = rust_type(ret) if ret else ("void", (), 0, False)
        out.append((name, r, params))
    return out


def rust_structs():
    src = strip_rust_comments(SYS.read_text())
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\]\s*(?:#\[[^\]]*\]\s*)*pub\s+struct\s+(\w+)\s*\{(.*?)\n\}", src,
                                 flags=re.S):
        fields = []
        for line in split_args(" ".join(body.split())):
            n, t = line.split(":", 1)
            n = n.replace("pub", "").strip()
            fields.append((n, t.strip()))
        out[name] = fields
    return out


def rust_consts():
    src = strip_rust_comments(SYS.read_text())
    return {k: int(v) for k, v in re.findall(r"pub\s+const\s+(VP8G_\w+)\s*:\s*\w+\s*=\s*(-?\d+)\s*;", src)}


# ---------------------------------------------------------------- tests

def test_every_declared_function_is_bound_in_order_with_equal_signatures():
    c = header_functions()
    r = rust_functions()
    assert len(c) >= 58
    assert [n for n, _, _ in r] == [n for n, _, _ in c]
    for (cn, cret, cpar), (rn, rret, rpar) in zip(c, r):
        assert rret == cret, (cn, "return", cret, rret)
        assert len(rpar) == len(cpar), (cn, "arity", len(cpar), len(rpar))
        for (pn, pt), (qn, qt, chain) in zip(cpar, rpar):
            assert qn == pn, (cn, "argument name", pn, qn)
            assert qt == pt, (cn, pn, pt, qt)
            if len(chain) > 1:  # T ** binds as *mut *mut T: outer levels mutable
                assert chain[:-1] == ["mut"] * (len(chain) - 1), (cn, pn, chain)


def test_constants_equal_the_header_defines():
    c, r = header_defines(), rust_consts()
    assert c, "no defines parsed"
    assert r == c


def rust_layout(structs, name):
    """repr(C) size, alignment and field offsets from the Rust field types."""
    def size_align(t):
        t = t.strip()
        if t.startswith("*"):
            return 8, 8
        m = re.match(r"^\[(.*);\s*(\d+)\]$", t)
        if m:
            s, a = size_align(m.group(1))
            return s * int(m.group(2)), a
        if t in structs:
            s, a, _ = rust_layout(structs, t)
            return s, a
        s = PRIM[C_OF_RUST[t]]
        return s, s

    off, align, offs = 0, 1, []
    for n, t in structs[name]:
        s, a = size_align(t)
        off = (off + a - 1) // a * a
        offs.append((n, off, s))
        off += s
        align = max(align, a)
    return (off + align - 1) // align * align, align, offs


def c_layout(structs):
    """sizeof / offsetof / field sizes from gcc on the header itself."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for s, fields in structs.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in fields:
            lines.append(f'printf("{s}.{f} %zu %zu\\n", offsetof({s}, {f}), sizeof((({s}*)0)->{f}));')
    lines += ["return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = Path(d) / "layout.c", Path(d) / "layout"
        src.write_text("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", str(src), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        k, *v = line.split()
        res[k] = tuple(int(x) for x in v)
    return res


def test_struct_fields_and_layouts_equal_the_c_compilers():
    cs, rs = header_structs(), rust_structs()
    opaque = header_opaque()
    for o in opaque:
        assert o in rs and rs[o][0][0] == "_private", f"opaque {o} not bound as an opaque struct"
    assert set(cs) <= set(rs), set(cs) - set(rs)
    lay = c_layout(cs)
    for s, fields in cs.items():
        rf = rs[s]
        assert [n for n, _ in rf] == [n for n, _ in fields], s
        for (n, ct), (_, rt) in zip(fields, rf):
            assert rust_type(rt) == ct, (s, n, ct, rt)
        size, _, offs = rust_layout(rs, s)
        assert size == lay[s][0], (s, "size", lay[s][0], size)
        for n, off, sz in offs:
            assert (off, sz) == lay[f"{s}.{n}"], (s, n, lay[f"{s}.{n}"], (off, sz))


def test_build_rs_compiles_what_build_py_compiles():
    import importlib.util
    spec = importlib.util.spec_from_file_location("vp8g_build", ROOT / "oxideav-vp8_amd" / "build.py")
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    rs = (RUST / "vp8g-sys" / "build.rs").read_text()
    m = re.search(r"const HIP_FLAGS: &\[&str\] = &\[(.*?)\];", rs, flags=re.S)
    flags = re.findall(r'"([^"]*)"', m.group(1))
    assert flags == b.HIP_FLAGS
    # build.rs takes every .hip / .cpp file of the source directory
    on_disk = sorted(p.name for p in b.CSRC.iterdir() if p.suffix in (".hip", ".cpp"))
    assert on_disk == sorted(b.SOURCES)
    assert "../../csrc" in rs and "-shared" in rs


def test_safe_crate_calls_only_bound_items():
    bound = {n for n, _, _ in rust_functions()} | set(rust_consts()) | set(rust_structs())
    used = set(re.findall(r"sys::(vp8g_\w+|VP8G_\w+)", SAFE.read_text()))
    assert used, "safe crate uses no bindings"
    assert used <= bound, used - bound
    for fn in ("build_arnr_altref", "build_arnr_altref_threaded", "predict_inter_mb", "mb_luma_sad_at_mv",
               "half_pixel_refine_luma", "quarter_pixel_refine_luma", "small_diamond_search_luma",
               "predict_split_mv", "reconstruct_inter_mb", "reconstruct_split_mv_mb", "luma_mad_batch"):
        assert re.search(rf"pub fn {fn}\b", SAFE.read_text()), fn


def test_library_exports_every_bound_function():
    lib = ROOT / "oxideav-vp8_amd" / "libvp8g.so"
    if not lib.exists():
        pytest.skip("libvp8g.so not built")
    nm = subprocess.run(["nm", "-D", "--defined-only", str(lib)], check=True, capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    missing = {n for n, _, _ in rust_functions()} - exported
    assert not missing, missing
