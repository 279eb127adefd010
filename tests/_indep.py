"""Independent numerical references used to pin the oracle (tests only).

Nothing here shares code with oracle/oracle.c: Gauss nodes come from scipy,
the triangle potential is the closed form of Wilton et al. (IEEE TAP 32, 1984)
/ Graglia (IEEE TAP 41, 1993), and the high-order cubature is a composite
Duffy-collapsed product rule.
"""
from __future__ import annotations

import numpy as np
from scipy.special import roots_legendre


def gl01(n):
    x, w = roots_legendre(n)
    return 0.5 * (x + 1.0), 0.5 * w


def collapsed(n):
    """n^2-point rule on the reference simplex {b1, b2 >= 0, b1 + b2 <= 1}; exact to degree 2n-2+."""
    x, w = gl01(n)
    s, t = np.meshgrid(x, x, indexing="ij")
    ws, wt = np.meshgrid(w, w, indexing="ij")
    b1 = s.ravel()
    b2 = (t * (1.0 - s)).ravel()
    ww = (ws * wt * (1.0 - s)).ravel()
    return b1, b2, ww


def subdivided(n, levels):
    """Composite rule: the reference simplex split into 4^levels similar
    sub-triangles, each with the collapsed n^2 rule."""
    tris = [np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])]
    for _ in range(levels):
        nt = []
        for T in tris:
            a, b, c = T
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nt += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = nt
    b1, b2, w = collapsed(n)
    P, W = [], []
    for T in tris:
        a, b, c = T
        # map reference (b1, b2) -> c + b1 (a - c) + b2 (b - c)
        pts = c[None, :] + b1[:, None] * (a - c)[None, :] + b2[:, None] * (b - c)[None, :]
        jac = abs((a - c)[0] * (b - c)[1] - (a - c)[1] * (b - c)[0])
        P.append(pts)
        W.append(w * jac)
    P = np.concatenate(P)
    return P[:, 0], P[:, 1], np.concatenate(W)


def cell_points(v, b1, b2):
    """x = v3 + b1 (v1 - v3) + b2 (v2 - v3); v is (3, 3) vertices."""
    return v[2][None, :] + b1[:, None] * (v[0] - v[2])[None, :] + b2[:, None] * (v[1] - v[2])[None, :]


def area(v):
    return 0.5 * np.linalg.norm(np.cross(v[0] - v[2], v[1] - v[2]))


def triangle_potential(v, r):
    """int_T 1/|r - y| dS_y for the flat triangle with vertices v (3, 3), at
    points r (m, 3).  Closed form of Wilton et al. 1984."""
    r = np.atleast_2d(r)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n = n / np.linalg.norm(n)
    d = (r - v[0]) @ n                    # signed height
    rho = r - d[:, None] * n[None, :]     # projection onto the plane
    ad = np.abs(d)
    out = np.zeros(r.shape[0])
    for i in range(3):
        a, b = v[i], v[(i + 1) % 3]       # edge from a (-) to b (+), counter-clockwise around n
        lv = (b - a) / np.linalg.norm(b - a)
        u = np.cross(lv, n)               # in-plane outward normal of the edge
        lp = (b - rho) @ lv
        lm = (a - rho) @ lv
        P0 = (a - rho) @ u                # signed distance of the projection to the edge line
        Rp = np.linalg.norm(r - b, axis=1)
        Rm = np.linalg.norm(r - a, axis=1)
        R0sq = P0 * P0 + d * d
        with np.errstate(divide="ignore", invalid="ignore"):
            logt = np.log((Rp + lp) / (Rm + lm))
            at = np.arctan2(P0 * lp, R0sq + ad * Rp) - np.arctan2(P0 * lm, R0sq + ad * Rm)
        term = P0 * logt - ad * at
        term = np.where(np.abs(P0) < 1e-14, 0.0, term)
        out += term
    return out


def split_rule(v, x, n, m=4):
    """Composite rule on the flat triangle v (3, 3) for integrands singular or
    nearly singular at the projection xp of the point x: the triangle is split
    into the 3m triangles (xp, a, b) over m pieces of each edge, each with
    the collapsed n^2 rule collapsed at xp (the Duffy transform removes a 1/R
    singularity there).  xp may lie outside v: the pieces then carry signed
    weights.  Returns points (k, 3) and weights (k,)."""
    nrm = np.cross(v[0] - v[2], v[1] - v[2])
    nrm = nrm / np.linalg.norm(nrm)
    xp = x - ((x - v[0]) @ nrm) * nrm
    b1, b2, w = collapsed(n)
    P, W = [], []
    for e in range(3):
        a0, a1 = v[e], v[(e + 1) % 3]
        for k in range(m):
            a = a0 + (a1 - a0) * k / m
            b = a0 + (a1 - a0) * (k + 1) / m
            sa = 0.5 * np.cross(a - xp, b - xp) @ nrm  # signed area
            if sa == 0.0:
                continue
            P.append(cell_points(np.array([xp, a, b]), b1, b2))
            W.append(w * 2.0 * sa)
    return np.concatenate(P), np.concatenate(W)


def barycentric(v, P):
    """Barycentric coordinates (b1, b2, 1 - b1 - b2) of points P (k, 3) in the
    plane of v, in the chart x = v3 + b1 (v1 - v3) + b2 (v2 - v3)."""
    M = np.stack([v[0] - v[2], v[1] - v[2]], axis=1)
    b = np.linalg.lstsq(M, (P - v[2]).T, rcond=None)[0].T
    return np.stack([b[:, 0], b[:, 1], 1.0 - b[:, 0] - b[:, 1]], axis=1)
