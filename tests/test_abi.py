"""CPU checks of the C ABI: the library loads and exports every symbol that
include/xr_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from qodeapplications_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "xr_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xr_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "xr_density" in syms and "xr_gemm_f64" in syms
    assert len(syms) >= 10


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libxrhip.so not built"
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_no_gpu_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        return
    try:
        _lib.Context(0)
    except RuntimeError as e:
        assert "no GPU" in str(e)
    else:
        raise AssertionError("context creation must fail without a GPU")


def test_conjugate_strings():
    from qodeapplications_amd.hermitian import conjugate_string
    pairs = {"a": "c", "aa": "cc", "caa": "cca", "ca": "ca", "ccaa": "ccaa", "caaa": "ccca", "ccaaa": "cccaa"}
    for op, adj in pairs.items():
        assert conjugate_string(op) == adj
        assert conjugate_string(adj) == op


def test_pair_job_struct_matches_the_header(tmp_path):
    """ctypes' PairJob has the size and field offsets the C compiler gives
    struct xr_pair_job (gcc on include/xr_hip.h)."""
    import subprocess
    fields = [f for f, _ in _lib.PairJob._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"xr_hip.h\"\nint main(void){\n"
                   "printf(\"%zu\\n\", sizeof(xr_pair_job));\n"
                   + "".join(f"printf(\"%zu\\n\", offsetof(xr_pair_job, {f}));\n" for f in fields)
                   + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(_lib.PairJob)
    assert vals[1:] == [getattr(_lib.PairJob, f).offset for f in fields]
