"""The C-ABI library loads and exports every symbol include/scft.h declares
(no device needed: nothing is called)."""
import ctypes as C
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "scft_amd", "libscft_hip.so")
HEADER = os.path.join(ROOT, "include", "scft.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)  # comments
    text = re.sub(r"^\s*#.*$", " ", text, flags=re.M)   # preprocessor
    text = re.sub(r"typedef\s+struct\s+\w+\s*\{.*?\}\s*\w+\s*;", " ", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", " ", text)
    text = text.replace('extern "C" {', " ")
    funcs, data = set(), set()
    for stmt in text.split(";"):
        s = " ".join(stmt.split())
        if not s or s == "}":
            continue
        m = re.match(r"^(?:extern\s+)?[\w\s\*]*?\b(\w+)\s*\(", s)
        if m:
            funcs.add(m.group(1))
            continue
        m = re.match(r"^extern\s+[\w\s\*]+?\b(\w+)$", s)
        if m:
            data.add(m.group(1))
    return funcs, data


def _lib():
    if not os.path.exists(LIB):
        subprocess.run([sys.executable, os.path.join(ROOT, "scft_amd", "build.py")], check=True)
    return C.CDLL(LIB, mode=C.RTLD_GLOBAL)


def test_header_parses_to_the_expected_entry_points():
    funcs, data = declared_symbols()
    for name in ("simple_FEM_1D_transient", "scft_init", "scft_batch_create",
                 "scft_batch_solve_resident", "scft_fixed_point", "scft_adm_chen_f"):
        assert name in funcs
    assert data == {"f0_given", "local_iteration", "funcerr"}


def test_library_exports_every_declared_symbol():
    lib = _lib()
    funcs, data = declared_symbols()
    missing = [n for n in sorted(funcs | data) if not hasattr(lib, n)]
    assert not missing, f"declared in scft.h but not exported: {missing}"


def test_binding_options_match_the_library_defaults():
    """scft_default_anderson (host only) fills the binding's AndersonOpts
    field by field as scft_amd.anderson_opts() does: the ctypes layout of
    scft_anderson_opts agrees with the header, its last field included."""
    _lib()
    import scft_amd as S
    o = S.AndersonOpts()
    o.inner_sweeps = -7  # overwritten by the library
    S.lib.scft_default_anderson(C.byref(o))
    ref = S.anderson_opts()
    for name, _ in S.AndersonOpts._fields_:
        assert getattr(o, name) == getattr(ref, name), name
    assert o.inner_sweeps == 1 and o.gs_shift_below == 0.0 and o.history == 10
