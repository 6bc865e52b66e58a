"""The C ABI boundary on the CPU: libcpolyg_hip.so loads, exports every function that
include/cpolyg.h declares, and rejects bad arguments before touching the device."""
import ctypes as C
import os
import subprocess

import pytest

from cpolyg_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = hip.lib()
    declared = hip.declared_symbols()
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_status_strings_and_version():
    lib = hip.lib()
    assert lib.cpg_version() >= 100
    assert lib.cpg_status_string(0) == b"ok"
    assert lib.cpg_status_string(4) == b"singular block"
    assert lib.cpg_status_string(6) == b"not converged"


def test_argument_validation_without_device():
    lib = hip.lib()
    rowptr = (C.c_int * 3)(0, 1, 3)
    colind = (C.c_int * 3)(0, 0, 5)   # 5 is out of range for 2 rows
    h = C.c_void_p()
    assert lib.cpg_bsr_create(2, 4, 3, rowptr, colind, C.byref(h)) == 1
    assert b"out of range" in lib.cpg_last_error()
    bad_rp = (C.c_int * 3)(0, 2, 1)
    assert lib.cpg_bsr_create(2, 4, 1, bad_rp, colind, C.byref(h)) == 1
    assert lib.cpg_bsr_create(2, 4, 3, rowptr, colind, None) == 1
    assert lib.cpg_krylov_create(10, 4, 0, 0, C.byref(h)) == 1
    assert lib.cpg_krylov_create(10, 4, 5, 9, C.byref(h)) == 1
    assert lib.cpg_bsr_matvec(None, None, None, None) == 1
    assert lib.cpg_prec_apply(None, None, None) == 1
    assert lib.cpg_gmres(None, None, None, None, None, None, None, None) == 1
    p = hip.EqParams()
    p.kind = 7
    assert lib.cpg_equation_create(None, C.byref(p), 3, C.byref(h)) == 1


def test_headers_compile_standalone(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "cpolyg.h"\nint main(void){cpg_eq_params p; (void)p; return 0;}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                        "-o", str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    src2 = tmp_path / "t.cpp"
    src2.write_text('#include "cpolyg.hpp"\nint main(){ cpolyg::hip::NoPreconditioner p; (void)p; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src2),
                        "-o", str(tmp_path / "t2.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_product_path_does_not_import_the_oracle():
    import sys
    for mod in list(sys.modules):
        if mod == "oracle" or mod.startswith("oracle."):
            pass
    code = ("import sys; sys.path.insert(0, %r); import cpolyg_amd, cpolyg_amd.hip, cpolyg_amd.mesh, "
            "cpolyg_amd.setups; print('oracle' in sys.modules)") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert out.stdout.strip() == "False", out.stderr
    # the HIP library does not link the oracle either
    deps = subprocess.run(["ldd", hip.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in deps
