"""The interior build's per-point tables, checked on the CPU.

build3d_interior_kernel (cfdv_amd/csrc/build.hip) evaluates the three flux
Jacobian directions on one code path -- direction 0's formulas
(NavierStokes3D.h:74-343) on the state with axes 0 and I swapped -- and
contracts the c_ij blocks (NavierStokes3D.h:347-477) against a table of
their distinct coefficients.  tests/host/point_tables_check.hip compiles
those __host__ __device__ functions for the host and compares them, on
random Gauss-point states, with the per-direction formulas calc_a3_dir<I>,
calc_b3_dir<I> and with accum_c3 (fdv_device.h), which the per-lane build
and the GPU parity tests pin against the oracle.  No GPU is used.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_point_tables_match_the_per_direction_formulas(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "point_tables_check")
    src = os.path.join(HERE, "host", "point_tables_check.hip")
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-Wno-unused-result", src, "-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_quad4_cosh_sinh_equal_the_host_libm(tmp_path):
    """The quad4 numerical diffusion's coth (FDVGalerkin.h:46-50) on the GPU
    uses fdlibm_expm1 / libm_sinh / libm_cosh (fdv_device.h); compiled for
    the host they equal this host's libm bit for bit on 4 million arguments,
    so the device blocks carry the same last bits as the oracle's."""
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "libm_coth_check")
    src = os.path.join(HERE, "host", "libm_coth_check.hip")
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-Wno-unused-result", src, "-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
