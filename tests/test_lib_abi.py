"""The C-ABI library without a GPU: it loads, exports every symbol that
include/imsrg_hip.h declares with the declared arity, rejects bad arguments
with an error code and message, and the Python package refuses to run
without it (no CPU fallback, no use of the oracle)."""
import ctypes as C
import os
import re
import shutil
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "imsrg_hip.h")
PKG = os.path.join(ROOT, "eom-imsrg_amd", "eom_imsrg")


def declared():
    """{name: number of parameters} of the extern "C" functions in the header"""
    text = re.sub(r"//[^\n]*", "", open(HEADER).read())
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(imsrg_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        args = " ".join(m.group(2).split())
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_header_parses():
    d = declared()
    assert len(d) >= 35
    for name in ("imsrg_commutator_110", "imsrg_commutator_222_ph", "imsrg_calculate_cross_coupled",
                 "imsrg_bch_expand", "imsrg_magnus_expand", "imsrg_euler_step", "imsrg_magnus_decouple"):
        assert name in d


def test_library_exports_every_declared_symbol():
    from eom_imsrg import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    nm = shutil.which("nm")
    if nm:
        out = subprocess.run([nm, "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
        exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
        missing = set(declared()) - exported
        assert not missing, missing


def test_bindings_match_header():
    from eom_imsrg import _lib
    d = declared()
    assert set(_lib.SIGNATURES) == set(d)
    for name, (_, args) in _lib.SIGNATURES.items():
        assert len(args) == d[name], name


def test_error_paths_without_gpu():
    from eom_imsrg import _lib
    lib = _lib.lib
    ia = lambda v: np.ascontiguousarray(v, np.int32).ctypes.data_as(C.POINTER(C.c_int))
    h = C.c_void_p()
    n, l, jj, tz, hole = [0, 0], [0, 0], [1, 2], [-1, 1], [1, 0]     # 2j = 2 is not odd
    arrs = [np.ascontiguousarray(x, np.int32) for x in (n, l, jj, tz, hole)]
    rc = lib.imsrg_basis_create(2, *[a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs], C.byref(h))
    assert rc == _lib.IMSRG_EINVAL and b"2j" in lib.imsrg_last_error()
    rc = lib.imsrg_basis_create(2, None, ia(l), ia(jj), ia(tz), ia(hole), C.byref(h))
    assert rc == _lib.IMSRG_EINVAL
    from conftest import basis
    b = basis(2)
    assert lib.imsrg_set_path(b.handle, 7) == _lib.IMSRG_EINVAL
    assert lib.imsrg_set_gemm_tile(b.handle, 9, 0) == _lib.IMSRG_EINVAL
    assert lib.imsrg_basis_pairs(b.handle, b.nblocks, 0, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_flow_destroy(None) == 0                       # like free(NULL)
    assert lib.imsrg_flow_set_graphs(None, 1) == _lib.IMSRG_EINVAL
    assert lib.imsrg_basis_size(None, 0) == -1
    # operator entry points check their arguments before touching the device
    x = C.c_void_p(16)                                             # never dereferenced: the checks fail first
    d = C.c_double()
    assert lib.imsrg_commutator_111(b.handle, x, 1, None, 1, x, -1, None) == _lib.IMSRG_EINVAL
    assert b"null" in lib.imsrg_last_error()
    assert lib.imsrg_commutator_122(b.handle, x, 1, x, 2, x, -1, None) == _lib.IMSRG_EINVAL
    assert b"herm" in lib.imsrg_last_error()
    assert lib.imsrg_commutator_222_pp_hh(b.handle, x, 1, x, -1, x, 0, x, x, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_commutator_220(b.handle, x, 1, x, 1, None, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_calculate_cross_coupled(b.handle, x, 3, x, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_mat_frob_norm(b.handle, None, C.byref(d), None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_build_gs_atan(b.handle, x, None, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_repackage(b.handle, x, 0, x, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_read_op_bin(b.handle, None, 1, x, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_commutator_111(None, x, 1, x, 1, x, -1, None) == _lib.IMSRG_EINVAL
    # the Python wrapper raises with the library's message
    try:
        _lib.check(lib.imsrg_set_path(b.handle, 7))
    except _lib.ImsrgError as e:
        assert "path" in str(e)
    else:
        raise AssertionError("no exception")


def test_package_refuses_to_run_without_library(tmp_path):
    dst = tmp_path / "eom_imsrg"
    dst.mkdir()
    for f in os.listdir(PKG):
        if f.endswith(".py"):
            shutil.copy(os.path.join(PKG, f), dst / f)
    r = subprocess.run([sys.executable, "-c", "import eom_imsrg"], cwd=str(tmp_path), capture_output=True, text=True,
                       env={**os.environ, "PYTHONPATH": str(tmp_path)})
    assert r.returncode != 0
    assert "no CPU fallback" in r.stderr


def test_product_does_not_use_the_oracle():
    pat = re.compile(r"(from|import)\s+oracle|liboracle|\borc_\w+")
    for d, _, files in os.walk(os.path.join(ROOT, "eom-imsrg_amd")):
        if "_build" in d or "__pycache__" in d:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h", ".f90")):
                txt = open(os.path.join(d, f)).read()
                code = "\n".join(ln for ln in txt.splitlines() if not ln.lstrip().startswith(("#", "//", "!")))
                if f == "build.py":
                    continue            # the build recipe also compiles the test oracle
                assert not pat.search(code), f"{f} refers to the oracle"
