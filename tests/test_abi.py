"""CPU tests of the C ABI boundary: the library loads and exports every symbol the
header declares (no device calls)."""
import ctypes
import os
import re

from athena_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "athena_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ath_(?:gpu|ml)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(gpu.LIBPATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(gpu.EXPORTS)


def test_python_mirror_struct_matches_header_order():
    text = open(os.path.join(ROOT, "include", "athena_gpu.h")).read()
    body = text[text.index("typedef struct ath_mesh_desc"):text.index("} ath_mesh_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b(\w+)(?:\[\d+\])?\s*[,;]", body)
    fields = [f[0] for f in gpu.MeshDescC._fields_]
    assert names == fields, (names, fields)
