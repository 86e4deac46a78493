"""cfdv_refine_flags_noalloc (no device allocation; the library's fixed
per-GPU scratch array) against cfdv_refine_flags, through the C ABI: flags
and s1 bit for bit, the flagged and tie counts equal, on two calls in a row
on one GPU -- a 16x8x4 ramp (512 elements), then a smaller 5x4x3 box (60
elements) -- so the second call runs on what the first left in the array.
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.fdv import FDVContext

pytestmark = pytest.mark.gpu


def _case(ctx):
    f, s1, ties = ctx.refine_flags()
    fn, s1n, tiesn, nf = ctx.refine_flags_noalloc()
    return f, s1, ties, fn, s1n, tiesn, nf


def test_refine_flags_noalloc_equals_refine_flags(gpu_available):
    meshes = [M.compression_ramp(16, 8, 4), M.mixed_bc_box(5, 4, 3)]
    for k, (m, th, U) in enumerate(meshes):
        assert m.n_elem <= 3700000
        ctx = FDVContext(m, th, U)
        f, s1, ties, fn, s1n, tiesn, nf = _case(ctx)
        print(f"call {k}: {m.n_elem} elements, {int(f.sum())} flagged, ties {ties}; "
              f"noalloc flagged {nf}, ties {tiesn}", flush=True)
        assert np.array_equal(fn, f)
        assert np.array_equal(s1n, s1)
        assert tiesn == ties and nf == int(f.sum())
        # without s1 the flags are the same
        fo, _, to, _ = ctx.refine_flags_noalloc(want_s1=False)
        assert np.array_equal(fo, f) and to == ties
