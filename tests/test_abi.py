"""CPU tests of the C ABI boundary: the library loads and exports every symbol the
header declares (no device calls)."""
import ctypes
import os
import re
import subprocess

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


def build_caller(outdir, name="abi_caller"):
    """Compile and link tests/<name>.cpp with g++ against the header and the library
    (host code only: hipMalloc and friends come from the HIP runtime library)."""
    exe = os.path.join(outdir, name)
    libdir = os.path.dirname(gpu.LIBPATH)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", name + ".cpp"), "-o", exe,
           "-L" + libdir, "-lathena_gpu", "-Wl,-rpath," + libdir,
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    return exe


def test_cpp_caller_compiles_and_links(tmp_path):
    # the header is valid C++ for an outside caller and every entry it uses links
    exe = build_caller(str(tmp_path))
    assert os.path.exists(exe)
    exe = build_caller(str(tmp_path), "abi_caller_smr")
    assert os.path.exists(exe)


def test_python_block_desc_matches_header_order():
    text = open(os.path.join(ROOT, "include", "athena_gpu.h")).read()
    body = text[text.index("typedef struct ath_block_desc"):text.index("} ath_block_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\*?(\w+)(?:\[\d+\])?\s*[,;]", body)
    fields = [f[0] for f in gpu.BlockDescC._fields_]
    assert names == fields, (names, fields)
    body = text[text.index("typedef struct ath_block_regs"):text.index("} ath_block_regs;")]
    names = re.findall(r"\*?(\w+)(?:\[\d+\])?\s*[,;]", body)
    assert names == [f[0] for f in gpu.BlockRegsC._fields_], names
