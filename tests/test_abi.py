"""The C-ABI library loads and exports every function include/zsc.h declares
(no GPU calls)."""
import ctypes
import os
import re

from lbm_bench_gpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "zsc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zsc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_kernel_names():
    names = set(declared_functions())
    for k in ("init_f", "init_g", "update_phi", "update_rho", "update_velocity", "collision_f",
              "collision_g", "stream_f", "stream_g"):
        assert f"zsc_{k}" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_struct_layouts_match_header():
    text = open(os.path.join(ROOT, "include", "zsc.h")).read()

    def fields(struct):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        out = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            typ, names = decl.split(None, 1)
            out += [(n.strip(), typ) for n in names.split(",")]
        return out

    for struct, cls in (("zsc_inputs", _lib.ZscInputs), ("zsc_params", _lib.ZscParams),
                        ("zsc_stats", _lib.ZscStats)):
        want = fields(struct)
        got = [(n, "int" if t is ctypes.c_int else "double") for n, t in cls._fields_]
        assert want == got, struct


def test_library_errors_without_a_context():
    L = _lib.lib()
    assert L.zsc_step(None, 1) == _lib.ZSC_ERR_ARG
    assert b"null context" in L.zsc_last_error()
    assert L.zsc_create(0, 4, 4, 0, None, ctypes.byref(ctypes.c_void_p())) == _lib.ZSC_ERR_ARG
    for bad in (-1, 5, 9):  # ZSC_F64 .. ZSC_F32D are 0 .. 4
        assert L.zsc_create(4, 4, 4, bad, None, ctypes.byref(ctypes.c_void_p())) == _lib.ZSC_ERR_ARG
    assert L.zsc_bytes_per_cell(7) == 0.0


def test_makefile_tracks_every_header():
    """Every header of csrc/ is a dependency of the objects (a header left out
    of HDRS leaves a stale library after an edit to it)."""
    csrc = os.path.join(ROOT, "lbm_bench_gpu_amd", "csrc")
    mk = open(os.path.join(csrc, "Makefile")).read()
    hdrs = re.search(r"^HDRS\s*:=\s*(.*)$", mk, re.M).group(1).split()
    for h in sorted(f for f in os.listdir(csrc) if f.endswith(".h")):
        assert h in hdrs, h
    assert "../../include/zsc.h" in hdrs
