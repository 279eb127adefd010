"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py)."""
import glob
import os

import numpy as np
import pytest

from beastgpu_amd import meshes, spaces
from beastgpu_amd.operators import Operator
from oracle import oracle as O

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    d = np.load(path)
    m = meshes.Mesh(d["vertices"], d["faces"])
    X = spaces.Space(m, int(d["kind"]), d["idx"], d["coeff"], int(d["nfun"]))
    op = Operator(int(d["op_kind"]), complex(d["alpha"]), complex(d["beta"]), complex(d["gamma"]))
    return op, X, d["Z"], d["counts"]


def test_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_fixture(path):
    op, X, Zg, counts = load(path)
    Z, c = O.assemble(op, X, X)
    assert np.array_equal(c, counts)
    assert np.linalg.norm(Z - Zg) <= 1e-14 * np.linalg.norm(Zg)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_gpu_matches_fixture(gpu_ctx, path):
    from beastgpu_amd.api import Assembler, GPUConfiguration

    op, X, Zg, counts = load(path)
    print("fixture", os.path.basename(path), flush=True)
    a = Assembler(op, X, X, GPUConfiguration(), ctx=gpu_ctx)
    Z = a.assemble().to_host()
    st = a.stats()
    a.close()
    assert [st["pairs"][k] for k in ("far", "common_vertex", "common_edge", "common_face", "near")] == list(counts)
    assert np.linalg.norm(Z - Zg) <= 1e-12 * np.linalg.norm(Zg)
