"""The allocation-free apply (bmi_matrix_apply_noalloc) against bmi_matrix_apply,
through the C ABI: two calls in a row, the second with a smaller matrix and
fewer right-hand sides; sizes beyond its scratch are refused."""
import numpy as np
import pytest

from beastgpu_amd import Maxwell3D, _bmi, meshes, spaces
from beastgpu_amd.api import Assembler, DeviceMatrix, GPUConfiguration

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


def _noalloc(M, x):
    f = _bmi.lib().bmi_matrix_apply_noalloc
    xf = np.asfortranarray(np.asarray(x, dtype=np.complex128).reshape(M.cols, -1))
    y = np.empty((M.rows, xf.shape[1]), dtype=np.complex128, order="F")
    _bmi.check(f(M.h, xf.ctypes.data, y.ctypes.data, xf.shape[1]))
    return y


def test_noalloc_apply_equals_apply_on_two_calls(gpu_ctx):
    rng = np.random.default_rng(8)
    for level, nrhs, output in ((4, 8, "c64"), (3, 3, "c32")):
        m = meshes.perturb_radial(meshes.icosphere(level), 30 + level, 0.03)
        X = spaces.raviartthomas(m)
        print(f"case noalloc L{level} nrhs={nrhs} {output}", flush=True)
        a = Assembler(Maxwell3D.singlelayer(wavenumber=2.2), X, X, GPUConfiguration(output=output), ctx=gpu_ctx)
        M = a.assemble()
        x = rng.standard_normal((X.nfun, nrhs)) + 1j * rng.standard_normal((X.nfun, nrhs))
        y_ref = M.apply(x)
        y = _noalloc(M, x)
        a.close()
        # same kernels and summation order as bmi_matrix_apply
        assert np.array_equal(y, y_ref), (level, nrhs)


def test_noalloc_apply_refuses_sizes_beyond_its_scratch(gpu_ctx):
    # 1 x 1,200,000 with 8 right-hand sides: x alone is 153.6 MB
    M = DeviceMatrix(gpu_ctx, 1, 1_200_000)
    x = np.ones((M.cols, 8), dtype=np.complex128)
    with pytest.raises(_bmi.BMIError) as e:
        _noalloc(M, x)
    assert e.value.status == _bmi.EINVAL and b"scratch" in _bmi.lib().bmi_last_error()
    M.close()
