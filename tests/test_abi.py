"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point that include/mipt_hip.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mipt_hip.h")
LIB = os.path.join(ROOT, "mipt_amd", "libmipt_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mipt_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_drop_ins():
    syms = declared_symbols()
    for name in ["mipt_hip_entropy_svd_f32", "mipt_hip_entropy_svd_f64", "mipt_hip_renyi2_f32", "mipt_hip_renyi2_f64",
                 "mipt_hip_rdm_subsystem_f32", "mipt_hip_rdm_subsystem_f64", "mipt_hip_measure_collapse_f32",
                 "mipt_hip_measure_collapse_f64", "mipt_hip_measure_collapse_last_error",
                 "mipt_hip_rho1_complex_f32", "mipt_hip_rho1_complex_f64"]:
        assert name in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmipt_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmipt_hip.so not built")
def test_argument_errors_need_no_gpu():
    import mipt_amd
    L = mipt_amd.lib()
    o = ctypes.c_int()
    p = ctypes.c_double()
    # invalid qubit / uniform are rejected before any device work
    assert L.mipt_hip_measure_collapse_f32(None, 4, 0, 0.5, ctypes.byref(o), ctypes.byref(p)) == -1
    assert b"Invalid" in L.mipt_hip_measure_collapse_last_error()
    assert L.mipt_hip_engine_create(0, None, None) == -1
    assert L.mipt_hip_batch_setup(None, None, 0) == -1
    assert mipt_amd.lib().mipt_hip_version().startswith(b"mipt_hip")


def test_header_is_plain_c(tmp_path):
    # the boundary is a C ABI: the header compiles as C99 (and C++17)
    import shutil
    import subprocess
    src = tmp_path / "t.c"
    src.write_text('#include "mipt_hip.h"\nint main(void) { return sizeof(mipt_hip_traj_spec) > 0 ? 0 : 1; }\n')
    inc = os.path.join(ROOT, "include")
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++17")):
        if shutil.which(cc) is None:
            pytest.skip(f"{cc} not available")
        r = subprocess.run([cc, std, "-Wall", "-Werror", "-fsyntax-only", "-I" + inc] +
                           (["-x", "c++"] if cc == "g++" else []) + [str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
