"""Writes the golden fixtures of tests/golden/ from the CPU FP64 oracle.

The reference's own golden matrix (cashed_results/matrix_ref_1.bin, read by
test/correctness.jl:25-30) is not in the reference repository, so these
fixtures come from the oracle, whose parts are pinned by tests/test_oracle.py.
Each fixture holds its inputs (mesh, assembly data, operator) and the matrix.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from beastgpu_amd import Helmholtz3D, Maxwell3D, meshes, spaces  # noqa: E402
from oracle import oracle as O  # noqa: E402


def cases():
    # the reference's correctness case: meshcuboid(1,1,1,0.5), lagrangec0d1, HH3D k=1 (test/correctness.jl:21-24)
    m = meshes.meshcuboid(1.0, 1.0, 1.0, 0.5)
    yield "hh3d_cube_d1_k1", Helmholtz3D.singlelayer(wavenumber=1.0), spaces.lagrangec0d1(m)
    m = meshes.icosphere(1)
    yield "efie_ico1_k1", Maxwell3D.singlelayer(wavenumber=1.0), spaces.raviartthomas(m)
    m = meshes.perturb_radial(meshes.icosphere(1), 1, 0.03)
    yield "hh3d_ico1pert_k2", Helmholtz3D.singlelayer(wavenumber=2.0), spaces.lagrangec0d1(m)


def main():
    for name, op, X in cases():
        Z, counts = O.assemble(op, X, X)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            vertices=X.mesh.vertices, faces=X.mesh.faces, kind=X.kind, idx=X.idx, coeff=X.coeff, nfun=X.nfun,
            op_kind=op.kind, alpha=op.alpha, beta=op.beta, gamma=op.gamma, Z=Z, counts=counts,
            ss_order=np.array([5, 5, 5]), far_orders=np.array([2, 3]))
        print(name, Z.shape, counts)


if __name__ == "__main__":
    main()
