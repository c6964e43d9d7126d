"""The Rust binding (sdr_amd/rust/sdr-amd) against include/sdr_amd.h, statically: no Rust
toolchain is installed here, so the crate is checked by reading it.

  * every function of the header is declared in src/sys.rs with the same parameter
    names, in order, and the Rust spelling of each C type, and nothing else is;
  * every struct of the header is a #[repr(C)] struct of sys.rs with the same fields in
    the same order and types (so the layouts agree), and every #define constant has
    the same value;
  * the safe wrapper calls only declared functions, and the shared library exports
    every one of them."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CRATE = ROOT / "sdr_amd" / "rust" / "sdr-amd"
HEADER = (ROOT / "include" / "sdr_amd.h").read_text()
SYS = (CRATE / "src" / "sys.rs").read_text() if (CRATE / "src" / "sys.rs").exists() else ""
LIB = (CRATE / "src" / "lib.rs").read_text() if (CRATE / "src" / "lib.rs").exists() else ""
pytestmark = pytest.mark.skipif(not SYS or not LIB, reason="crate sources not present")

C_TO_RUST = {"double": "f64", "uint64_t": "u64", "int32_t": "i32", "uint32_t": "u32", "uint8_t": "u8",
             "size_t": "usize", "char": "c_char", "void": "c_void"}


def rust_type(ct: str) -> str:
    ct = " ".join(ct.split())
    stars = ct.count("*")
    base = ct.replace("*", "").replace("const", "").strip()
    base = C_TO_RUST.get(base, base)
    if stars == 0:
        return base
    if stars == 1:
        return ("*const " if ct.startswith("const") else "*mut ") + base
    assert stars == 2
    return "*mut *mut " + base


def header_functions():
    body = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"\n((?:const )?[a-z_0-9]+\*?)\s+(sdr_[a-z_0-9]+)\(([^;]*?)\);", body, flags=re.S):
        params = []
        args = " ".join(args.split())
        if args != "void":
            for p in args.split(","):
                m = re.match(r"(.*?)\s*(\w+)$", p.strip())
                params.append((m.group(2), rust_type(m.group(1))))
        out[name] = (ret, params)
    return out


def header_structs():
    body = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    out = {}
    for name, fields in re.findall(r"typedef struct (\w+) \{(.*?)\} \w+;", body, flags=re.S):
        fl = []
        for decl in fields.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"((?:const )?[a-z_0-9]+\*?)\s+(.*)", decl)
            for nm in m.group(2).split(","):
                fl.append((nm.strip(), rust_type(m.group(1))))
        out[name] = fl
    return out


def rust_functions():
    block = SYS[SYS.index('extern "C" {'):]
    out = {}
    for name, args, ret in re.findall(r"pub fn (sdr_\w+)\((.*?)\)( -> [^;]+)?;", block):
        params = []
        if args.strip():
            for p in args.split(","):
                nm, ty = p.split(":", 1)
                params.append((nm.strip(), " ".join(ty.split())))
        out[name] = (ret.strip(" ->") if ret else "", params)
    return out


def rust_structs():
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\]\n(?:#\[derive\([^)]*\)\]\n)?pub struct (\w+) \{(.*?)\n\}", SYS, flags=re.S):
        fl = []
        for line in body.strip().splitlines():
            line = line.strip().rstrip(",")
            if line.startswith("pub "):
                nm, ty = line[4:].split(":", 1)
                fl.append((nm.strip(), " ".join(ty.split())))
        out[name] = fl
    return out


def test_every_header_function_is_declared_alike():
    h, r = header_functions(), rust_functions()
    assert set(h) == set(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name, (ret, params) in h.items():
        rret, rparams = r[name]
        want = {"int": "c_int", "const char*": "*const c_char"}[ret]
        assert rret == want, name
        assert rparams == params, name


def test_every_header_struct_has_the_same_layout():
    h, r = header_structs(), rust_structs()
    assert set(h) <= set(r)
    for name, fields in h.items():
        assert r[name] == fields, name
    for opaque in ("sdr_solver", "sdr_metrics", "sdr_frame_writer"):
        assert r[opaque] == [], opaque  # zero-sized, only ever behind a pointer
        assert f"typedef struct {opaque} {opaque};" in HEADER


def test_constants_match():
    h = dict(re.findall(r"#define (SDR_[A-Z0-9_]+) \(?(-?\d+)u?(?: << \d+)?\)?", HEADER))
    r = dict(re.findall(r"pub const (SDR_[A-Z0-9_]+): \w+ = (-?\d+)", SYS))
    assert set(h) == set(r)
    assert h == r
    assert "pub const SDR_NOALLOC_SCRATCH_BYTES: u64 = 64 << 20;" in SYS


def test_wrapper_calls_only_declared_functions():
    declared = set(rust_functions())
    used = set(re.findall(r"sys::(sdr_\w+)\(", LIB))
    assert used and used <= declared, sorted(used - declared)
    # the reference API the wrapper mirrors
    for item in ("pub fn new(solid: &Sdf", "pub fn substep(&mut self, dt: f64)", "pub fn step(&mut self, dt: f64)",
                 "pub fn particles(&mut self)", "pub fn project_capturing_with_source(", "pub fn record(",
                 "pub fn summary(&self)", "pub fn write_points_vtk(", "pub fn read_points_vtk(",
                 "pub fn from_mesh("):
        assert item in LIB, item


def test_braces_balance():
    """A cheap syntax guard in place of a compiler: balanced delimiters outside strings."""
    for text in (SYS, LIB, (CRATE / "build.rs").read_text()):
        code = re.sub(r'"(?:\\.|[^"\\])*"', '""', re.sub(r"//[^\n]*", "", text))
        for a, b in ("{}", "()", "[]"):
            assert code.count(a) == code.count(b), (a, b)


def test_library_exports_every_declared_function():
    so = ROOT / "sdr_amd" / "libsdr_amd.so"
    if not so.exists():
        pytest.skip("library not built")
    dll = C.CDLL(str(so))
    for name in rust_functions():
        assert hasattr(dll, name), name
