"""Basis-function spaces on triangle meshes, as assembly data.

A space maps each local shape i of each cell K to one global function index
and a coefficient -- the (index, coefficient) pairs that BEAST's
``assemblydata`` returns and that the reference flattens to (3, N) arrays of
the first slot (load_parameters.jl:50-55).  Here they are stored as (nt, 3)
arrays ``idx`` (int64, -1 = no function) and ``coeff`` (float64).

* ``lagrangec0d1`` -- continuous piecewise-linear (P1) functions, one per
  vertex; local shape i on a cell is its barycentric coordinate of vertex i.
* ``raviartthomas`` -- RWG functions, one per interior edge; local shape i is
  f_i(x) = (x - v_i) / (2 A) (unit flux through the edge opposite vertex i,
  divergence 1/A), with coefficient +1 on the first cell that lists the edge
  and -1 on the second, so the normal component is continuous.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .meshes import Mesh

P1 = 1
RWG = 2


@dataclasses.dataclass
class Space:
    mesh: Mesh
    kind: int
    idx: np.ndarray  # (nt, 3) int64
    coeff: np.ndarray  # (nt, 3) float64
    nfun: int

    def __post_init__(self):
        self.idx = np.ascontiguousarray(self.idx, dtype=np.int64)
        self.coeff = np.ascontiguousarray(self.coeff, dtype=np.float64)
        if self.idx.shape != (self.mesh.nt, 3) or self.coeff.shape != (self.mesh.nt, 3):
            raise ValueError("idx/coeff must be (nt, 3)")

    def numfunctions(self) -> int:
        return self.nfun


def lagrangec0d1(mesh: Mesh) -> Space:
    idx = mesh.faces.astype(np.int64)
    coeff = np.ones((mesh.nt, 3))
    return Space(mesh, P1, idx, coeff, mesh.nv)


def edges_of(mesh: Mesh):
    """Edge numbering in order of first appearance over (cell, local shape);
    local shape i of a cell is the edge opposite its vertex i.  Returns
    (edge_of_shape (nt,3), cells_of_edge list)."""
    f = mesh.faces
    a = np.stack([f[:, 1], f[:, 2], f[:, 0]], axis=1)
    b = np.stack([f[:, 2], f[:, 0], f[:, 1]], axis=1)
    lo = np.minimum(a, b).ravel().astype(np.int64)
    hi = np.maximum(a, b).ravel().astype(np.int64)
    key = lo * (mesh.nv + 1) + hi
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # renumber unique keys by first appearance
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    eid = rank[inv].reshape(mesh.nt, 3)
    return eid


def raviartthomas(mesh: Mesh) -> Space:
    eid = edges_of(mesh)
    ne = int(eid.max()) + 1
    counts = np.bincount(eid.ravel(), minlength=ne)
    if (counts > 2).any():
        raise ValueError("non-manifold edge: an RWG function needs at most two cells per edge")
    interior = counts == 2
    # renumber interior edges only, keeping first-appearance order
    new = -np.ones(ne, dtype=np.int64)
    new[interior] = np.arange(int(interior.sum()))
    idx = new[eid]
    coeff = np.zeros((mesh.nt, 3))
    seen = np.zeros(ne, dtype=bool)
    flat = eid.ravel()
    cf = coeff.ravel()
    for s in range(flat.size):
        e = flat[s]
        if not interior[e]:
            continue
        cf[s] = -1.0 if seen[e] else 1.0
        seen[e] = True
    coeff = cf.reshape(mesh.nt, 3)
    return Space(mesh, RWG, idx, coeff, int(interior.sum()))
