"""Host-side tests (no GPU): meshes, spaces, the C ABI surface of libbmi.so."""
import ctypes
import os
import re

import numpy as np
import pytest

from beastgpu_amd import _bmi, meshes, spaces

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ----------------------------------------------------------------- meshes

@pytest.mark.parametrize("level,nv,nt", [(0, 12, 20), (1, 42, 80), (3, 642, 1280), (5, 10242, 20480)])
def test_icosphere_counts_and_orientation(level, nv, nt):
    m = meshes.icosphere(level)
    assert (m.nv, m.nt) == (nv, nt)
    v = m.vertices[m.faces]
    n = np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2])
    assert (np.einsum("ij,ij->i", n, v.mean(axis=1)) > 0).all()
    assert np.allclose(np.linalg.norm(m.vertices, axis=1), 1.0)
    assert abs(m.areas().sum() - 4 * np.pi) < 0.5 * 4 ** -level * 4 * np.pi


def test_perturbation_amplitude_and_seed():
    m = meshes.icosphere(3)
    a = meshes.perturb_radial(m, 2, 0.03)
    b = meshes.perturb_radial(m, 2, 0.03)
    c = meshes.perturb_radial(m, 3, 0.03)
    r = np.linalg.norm(a.vertices, axis=1)
    assert abs(np.abs(r - 1.0).max() - 0.03) < 1e-12
    assert np.array_equal(a.vertices, b.vertices)
    assert not np.array_equal(a.vertices, c.vertices)


@pytest.mark.parametrize("d", [1, 2, 24])
def test_meshcuboid_degree_pattern(d):
    m = meshes.meshcuboid(1.0, 1.0, 1.0, 0.5 / d)
    n = 2 * d
    assert m.nt == 12 * n * n
    assert m.nv == 6 * n * n + 2
    deg = np.bincount(m.vertex_degrees())
    assert deg[3] == 2 and deg[5] == 6 and deg[6] == m.nv - 8
    if d == 24:
        # reference's recorded common-vertex count (data/GPU/Saut-Swab.txt:50-55)
        assert int((m.vertex_degrees().astype(np.int64) ** 2).sum() - 9 * m.nt) == 248784


def test_ellipsoid_60k():
    m = meshes.ellipsoid()
    assert m.nt == 60000
    X = spaces.raviartthomas(m)
    assert X.nfun == 90000
    assert (X.idx >= 0).all()
    r = (m.vertices / np.array([1.0, 0.7, 0.5])) ** 2
    assert np.allclose(r.sum(axis=1), 1.0)
    assert m.areas().min() > 0.2 * m.areas().mean()


# ----------------------------------------------------------------- spaces

def test_rwg_space_structure():
    m = meshes.icosphere(2)
    X = spaces.raviartthomas(m)
    assert X.nfun == 3 * m.nt // 2
    cnt_plus = np.bincount(X.idx[X.coeff > 0], minlength=X.nfun)
    cnt_minus = np.bincount(X.idx[X.coeff < 0], minlength=X.nfun)
    assert (cnt_plus == 1).all() and (cnt_minus == 1).all()
    # local shape i sits on the edge opposite vertex i
    for K in range(10):
        f = m.faces[K]
        for i in range(3):
            e = X.idx[K, i]
            other = np.argwhere(X.idx == e)
            K2, i2 = [tuple(o) for o in other if o[0] != K][0]
            assert set(m.faces[K2]) - {m.faces[K2][i2]} == set(f) - {f[i]}


def test_rwg_open_mesh_excludes_boundary_edges():
    m = meshes.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float),
                    np.array([[0, 1, 2], [1, 3, 2]], np.int32))
    X = spaces.raviartthomas(m)
    assert X.nfun == 1
    assert (X.idx >= 0).sum() == 2


def test_lagrange_space():
    m = meshes.icosphere(1)
    X = spaces.lagrangec0d1(m)
    assert X.nfun == m.nv
    assert np.array_equal(X.idx, m.faces)
    assert (X.coeff == 1.0).all()


# ------------------------------------------------------------------ C ABI

def _declared_entry_points():
    text = open(os.path.join(ROOT, "include", "bmi.h")).read()
    return sorted(set(re.findall(r"\b(bmi_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _bmi.lib()
    declared = _declared_entry_points()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_bmi.EXPORTS) == declared


def test_options_default_and_version():
    o = _bmi.default_options()
    assert (o.far_test_order, o.far_trial_order, list(o.ss_order), o.near_order) == (2, 3, [5, 5, 5], 4)
    assert o.precision == _bmi.PREC_FP64 and o.near_factor == 2.0
    assert b"gfx950" in _bmi.lib().bmi_version()


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_bmi.Operator) == 4 + 4 + 6 * 8  # int32 + padding + 3 complex
    assert ctypes.sizeof(_bmi.Options) == 6 * 4 + 4 + 4 + 8
    assert ctypes.sizeof(_bmi.Stats) == 4 * 8 + 6 * 8 + 3 * 4 + 4 + 8


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_fails_loudly():
    from beastgpu_amd.api import Context

    with pytest.raises(_bmi.BMIError) as e:
        Context(0)
    assert e.value.status in (_bmi.EINVAL, _bmi.EDEVICE)


def test_null_arguments_are_rejected():
    L = _bmi.lib()
    assert L.bmi_context_create(0, None) == _bmi.EINVAL
    assert L.bmi_assemble(None, None) == _bmi.EINVAL
    assert b"NULL" in L.bmi_last_error()
    assert L.bmi_matrix_free(None) == _bmi.OK


def _c_layout(struct: str, fields: list) -> list:
    """sizeof and offsetof of a struct from include/bmi.h, by compiling a probe with gcc."""
    import subprocess
    import tempfile

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bmi.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({struct}));']
    lines += [f'  printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        with open(src, "w") as fh:
            fh.write("\n".join(lines))
        subprocess.run(["gcc", "-I", os.path.join(root, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("cname,pyname", [
    ("bmi_operator", "Operator"), ("bmi_options", "Options"), ("bmi_stats", "Stats"),
    ("bmi_server_options", "ServerOptions"), ("bmi_request", "Request"), ("bmi_result", "Result"),
    ("bmi_batch_stats", "BatchStats"), ("bmi_rules", "Rules"),
])
def test_ctypes_layouts_match_c_compiler(cname, pyname):
    S = getattr(_bmi, pyname)
    names = [f[0] for f in S._fields_]
    want = _c_layout(cname, names)
    got = [ctypes.sizeof(S)] + [getattr(S, n).offset for n in names]
    assert got == want


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_server_without_device_fails_loudly():
    from beastgpu_amd.serving import Server

    with pytest.raises(_bmi.BMIError):
        Server(slots=2)


def _assembly_request(mesh, X, precision=_bmi.PREC_FP64, near_factor=2.0):
    keep = [np.ascontiguousarray(mesh.vertices, dtype=np.float64), np.ascontiguousarray(mesh.faces, dtype=np.int32),
            np.ascontiguousarray(X.idx, dtype=np.int64), np.ascontiguousarray(X.coeff, dtype=np.float64)]
    r = _bmi.Request()
    r.kind = _bmi.REQ_ASSEMBLE_APPLY
    r.xyz, r.nv = keep[0].ctypes.data, mesh.nv
    r.tri, r.nt = keep[1].ctypes.data, mesh.nt
    r.space_kind = _bmi.SPACE_RWG
    r.gidx, r.coeff, r.nfun = keep[2].ctypes.data, keep[3].ctypes.data, X.nfun
    r.opts = _bmi.default_options()
    r.opts.precision = precision
    r.opts.near_factor = near_factor
    r.dtype = _bmi.C64
    return r, keep


def _special_candidates(mesh, reach_mixed):
    """Brute force: cell pairs that can be special -- touching (a shared
    vertex), near (min vertex distance below 0.8 sqrt(largest area)) or, for
    the mixed path, with centroids closer than reach_mixed."""
    v = mesh.vertices[mesh.faces]
    c = v.mean(axis=1)
    a = mesh.areas()
    dmin = np.full((mesh.nt, mesh.nt), np.inf)
    for i in range(3):
        for j in range(3):
            d = np.linalg.norm(v[:, None, i, :] - v[None, :, j, :], axis=2)
            dmin = np.minimum(dmin, d)
    cd = np.linalg.norm(c[:, None, :] - c[None, :, :], axis=2)
    cand = (dmin < 1e-12) | (dmin < 0.8 * np.sqrt(a.max())) | (cd < reach_mixed)
    return int(cand.sum())


@pytest.mark.parametrize("mixed", [False, True])
def test_request_bytes_bound_covers_special_pairs(mixed):
    """The admission bound covers 3 entries of 154 bytes for every cell pair
    that can be special, counted by brute force (jittered ellipsoid: thin
    cells, many FP64-kept far pairs on the mixed path), and the matrix."""
    m = meshes.ellipsoid(cells=(16, 10, 10))
    X = spaces.raviartthomas(m)
    r, keep = _assembly_request(m, X, _bmi.PREC_MIXED if mixed else _bmi.PREC_FP64)
    mb, pb = ctypes.c_int64(), ctypes.c_int64()
    _bmi.check(_bmi.lib().bmi_request_bytes(ctypes.byref(r), ctypes.byref(mb), ctypes.byref(pb)))
    assert mb.value == X.nfun ** 2 * 16
    n = _special_candidates(m, 2.0 * m.max_edge() if mixed else 0.0)
    assert pb.value >= 3 * 154 * n
    assert pb.value <= (64 << 20) + m.nt * 8192 + (m.nt / 8 + 64) ** 2 * 8 + 3 * 154 * m.nt ** 2
    bad = np.full(3 * m.nt, m.nv + 5, dtype=np.int32)
    r.tri = bad.ctypes.data
    assert _bmi.lib().bmi_request_bytes(ctypes.byref(r), ctypes.byref(mb), ctypes.byref(pb)) == _bmi.EINVAL


def test_tile_autotuner_signature_and_refusal():
    """The tile autotuner's cache key separates what changes the fastest
    shape (operator, precision, space kind, size class) and no more; a tile
    that is neither a shape nor "auto" is refused before any device call."""
    from types import SimpleNamespace

    from beastgpu_amd import Helmholtz3D, Maxwell3D
    from beastgpu_amd.api import TILE_CANDIDATES, Assembler, GPUConfiguration, tile_signature

    ctx = SimpleNamespace(device=0)
    m2a = meshes.perturb_radial(meshes.icosphere(2), 1, 0.03)
    m2b = meshes.perturb_radial(meshes.icosphere(2), 2, 0.03)
    m3 = meshes.icosphere(3)
    Xa, Xb, X3 = spaces.raviartthomas(m2a), spaces.raviartthomas(m2b), spaces.raviartthomas(m3)
    efie, mfie = Maxwell3D.singlelayer(wavenumber=2.0), Maxwell3D.doublelayer(wavenumber=2.0)
    cfg = GPUConfiguration(tile="auto")
    sig = tile_signature(ctx, efie, Xa, Xa, cfg)
    assert sig == tile_signature(ctx, Maxwell3D.singlelayer(wavenumber=3.5), Xb, Xb, cfg)
    assert sig != tile_signature(ctx, mfie, Xa, Xa, cfg)
    assert sig != tile_signature(ctx, efie, X3, X3, cfg)
    assert sig != tile_signature(ctx, efie, Xa, Xa, GPUConfiguration(tile="auto", precision="mixed"))
    assert sig != tile_signature(ctx, efie, Xa, Xb, cfg)
    # caller tables: keyed by their point counts, not by object identity
    from beastgpu_amd.api import QuadratureRules
    gl = tuple((np.array([0.5]), np.array([1.0])) for _ in range(3))
    def rules(nt, nb):
        return QuadratureRules(np.full((nt, 2), 1 / 3), np.full(nt, 0.5 / nt), np.full((nb, 2), 1 / 3),
                               np.full(nb, 0.5 / nb), gl)
    r34 = tile_signature(ctx, efie, Xa, Xa, GPUConfiguration(tile="auto", rules=rules(3, 4)))
    assert r34 == tile_signature(ctx, efie, Xa, Xa, GPUConfiguration(tile="auto", rules=rules(3, 4)))
    assert r34 != tile_signature(ctx, efie, Xa, Xa, GPUConfiguration(tile="auto", rules=rules(4, 4)))
    P = spaces.lagrangec0d1(m2a)
    assert sig != tile_signature(ctx, Helmholtz3D.singlelayer(wavenumber=2.0), P, P, cfg)
    for kind, shapes in TILE_CANDIDATES.items():
        assert (0, 0) in shapes and len(set(shapes)) == len(shapes)
        lim = (64, 40) if kind == spaces.RWG else (128, 20)
        assert all(0 <= r <= lim[0] and (c == 0 or 3 <= c <= lim[1]) for r, c in shapes)
    with pytest.raises(ValueError):
        Assembler(efie, Xa, Xa, GPUConfiguration(tile="fastest"))
