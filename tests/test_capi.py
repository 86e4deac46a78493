"""The C-ABI library: it loads on a host without a GPU, exports every entry
point that include/cfdv.h declares, and rejects malformed arguments before
touching the device.  No compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cfdv_amd import _lib
from cfdv_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cfdv.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfdv_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from cfdv_amd import build_lib
        build_lib.build()
    return _lib.load()


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol(lib):
    for name in _declared():
        assert hasattr(lib, name), name


def test_header_compiles_as_c():
    import shutil
    import subprocess
    import tempfile
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.c")
        with open(src, "w") as fh:
            fh.write('#include "cfdv.h"\nint main(void){cfdv_ctx *c = 0; (void)c; return 0;}\n')
        subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-c", src, "-o", os.path.join(d, "t.o")], check=True)


def test_create_rejects_bad_arguments_before_device_use(lib):
    m, th, _ = M.mixed_bc_box(2, 2, 2)
    conn = np.ascontiguousarray(m.conn, dtype=np.int32)
    x = np.ascontiguousarray(m.x)
    bc = np.ascontiguousarray(m.bc, dtype=np.int32)
    d = np.ascontiguousarray(m.dirichlet, dtype=np.uint8)

    def mesh_struct(ndim=3, nnod=8, neqn=5, n_elem=m.n_elem):
        return _lib.Mesh(ndim, nnod, neqn, n_elem, m.n_node,
                         conn.ctypes.data_as(C.POINTER(C.c_int32)),
                         x.ctypes.data_as(C.POINTER(C.c_double)),
                         bc.ctypes.data_as(C.POINTER(C.c_int32)),
                         d.ctypes.data_as(C.POINTER(C.c_uint8)))

    t = _lib.ThermoS(th.gamma, th.mach, th.reynolds, th.prandtl, th.tinf, th.twall,
                     int(th.adiabatic))
    h = C.c_void_p()
    for ms, flags in [(mesh_struct(ndim=2), 1), (mesh_struct(nnod=4), 1),
                      (mesh_struct(n_elem=0), 1), (mesh_struct(), 0),
                      (mesh_struct(ndim=2, nnod=4, neqn=4), _lib.CFDV_BLOCKS_ROWS),  # rows alone: hex8 only
                      (mesh_struct(), _lib.CFDV_BLOCKS_FP32 | _lib.CFDV_BLOCKS_ROWS),
                      (mesh_struct(), 1 | 8)]:                                      # unknown flag
        assert lib.cfdv_create(C.byref(ms), C.byref(t), flags, C.byref(h)) == _lib.CFDV_ERR_ARG
        assert not h.value
    assert lib.cfdv_create(None, C.byref(t), 1, C.byref(h)) == _lib.CFDV_ERR_ARG
    assert lib.cfdv_last_error(None) == b"null context"
    assert lib.cfdv_mesh_info(None, None, None, None, None) == _lib.CFDV_ERR_ARG


def test_fdv_module_raises_without_library(monkeypatch, tmp_path):
    """The product path has no CPU fallback: a missing library is an error."""
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libcfdv_amd.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.load()
