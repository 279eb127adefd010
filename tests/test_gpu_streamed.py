"""Streamed assembly into host memory (bmi_assemble_to_host, the "cpu"
write-back strategy): row blocks of test panels through device staging
blocks under a memory budget give bitwise the matrix of the resident path."""
import ctypes

import numpy as np
import pytest

from beastgpu_amd import Helmholtz3D, Maxwell3D, _bmi, meshes, spaces
from beastgpu_amd.api import Assembler, GPUConfiguration, assemble_gpu
from beastgpu_amd.spaces import Space
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _resident(op, X, Y, cfg, ctx):
    a = Assembler(op, X, Y, cfg, ctx=ctx)
    Z = a.assemble().to_host()
    a.close()
    return Z


def test_icosphere_l5_under_a_4gb_budget_equals_resident(gpu_ctx):
    """The 30,720-RWG EFIE matrix (15.1 GB) through 4 GB of staging (8 blocks
    of 60 panels, two on the device at a time): bitwise the resident result."""
    m = meshes.perturb_radial(meshes.icosphere(5), seed=2, amplitude=0.03)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=2.0)
    print("case streamed icosphere L5, 4 GB", flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    Zr = a.assemble().to_host()
    a.close()
    s = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx, resident=False)
    assert s.matrix is None
    Zs = s.assemble_to_host(budget=4 << 30)
    st = s.stats()
    s.close()
    print("streamed stats", st, flush=True)
    assert st["far_launches"] == 6 * 8
    assert np.array_equal(Zs, Zr)


@pytest.mark.parametrize("case", ["efie-fp64", "efie-mixed-c32", "hh3d-p1", "mfie", "uncovered"])
def test_small_budgets_equal_resident(gpu_ctx, case):
    """Budgets of one or a few panels per block, wide (Lagrange) panels, C32,
    a space with functions that have no shape (zero rows and columns)."""
    m = meshes.perturb_radial(meshes.icosphere(3), seed=9, amplitude=0.02)
    cfg = GPUConfiguration()
    if case.startswith("efie"):
        op, X = Maxwell3D.singlelayer(wavenumber=2.5), spaces.raviartthomas(m)
        if case == "efie-mixed-c32":
            cfg = GPUConfiguration(precision="mixed", output="c32")
    elif case == "mfie":
        op, X = Maxwell3D.doublelayer(wavenumber=2.5), spaces.raviartthomas(m)
    elif case == "hh3d-p1":
        op, X = Helmholtz3D.singlelayer(wavenumber=1.5), spaces.lagrangec0d1(m)
    else:
        op, B = Maxwell3D.singlelayer(wavenumber=2.5), spaces.raviartthomas(m)
        X = Space(m, B.kind, B.idx, B.coeff, B.nfun + 70)  # 70 functions without shapes
    es = 16 if cfg.output == "c64" else 8
    prows = 128 if X.kind == spaces.P1 else 64
    print(f"case streamed {case}", flush=True)
    Zr = _resident(op, X, X, cfg, gpu_ctx)
    s = Assembler(op, X, X, cfg, ctx=gpu_ctx, resident=False)
    for panels in (1, 3):
        Zs = s.assemble_to_host(budget=2 * panels * prows * X.nfun * es)
        assert np.array_equal(Zs, Zr), panels
    Zs = s.assemble_to_host(budget=64 << 30)  # one block per eighth of the panels
    assert np.array_equal(Zs, Zr)
    with pytest.raises(_bmi.BMIError) as e:
        s.assemble_to_host(budget=2 * prows * X.nfun * es - 1)
    assert e.value.status == _bmi.ENOMEM
    s.close()
    if case == "uncovered":
        assert not Zr[-70:, :].any() and not Zr[:, -70:].any()


def test_assemble_gpu_both_strategies_match_oracle(gpu_ctx):
    """The drop-in entry point end to end: writeBackStrategy "gpu" and "cpu"
    give the same host matrix, within 1e-12 of the oracle."""
    m = meshes.perturb_radial(meshes.icosphere(3), seed=21, amplitude=0.03)
    X = spaces.raviartthomas(m)
    op = Maxwell3D.singlelayer(wavenumber=1.7)
    print("case assemble_gpu", flush=True)
    Zg = assemble_gpu(op, X, X, GPUConfiguration(writeBackStrategy="gpu"))
    Zc = assemble_gpu(op, X, X, GPUConfiguration(GPU_budget_pipeline_result=8 << 20), writeBackStrategy="cpu")
    assert Zg.flags.f_contiguous and Zg.dtype == np.complex128
    assert np.array_equal(Zg, Zc)
    Zref, _ = O.assemble(op, X, X)
    assert np.linalg.norm(Zg - Zref) <= 1e-12 * np.linalg.norm(Zref)
    with pytest.raises(ValueError):
        assemble_gpu(op, X, X, writeBackStrategy="host")
