"""Truncated prolate-ellipsoid ventricular shells (benchmark and test meshes).

A shell is the region between an inner ellipsoid (semi-axes b_in, b_in, a_in)
and an outer one (b_out, b_out, a_out), truncated at the base plane z = 0:

    P(s, theta, phi) = (b(s) sin theta cos phi, b(s) sin theta sin phi, -a(s) cos theta)
    a(s) = a_in + s (a_out - a_in),  b(s) = b_in + s (b_out - b_in),  s in [0, 1]

with theta from theta_min (a small apical opening, where the grid would
otherwise collapse to a point) to pi/2 (the base).  The (s, theta, phi) grid,
periodic in phi, is cut into hexahedra of edge about h and each hexahedron
into six tetrahedra around its main diagonal (the split ConstructRegularSlabMesh
uses, so neighbouring hexahedra share face diagonals, across the phi seam too).

Fibres lie in the wall's tangent plane and rotate linearly through the wall,
from -60 degrees at the endocardium to +60 at the epicardium, measured from the
circumferential direction towards the apex-to-base direction.
"""
import numpy as np

# the six tetrahedra of a hexahedron with corners numbered by bits (s, theta, phi)
# = (1, 2, 4): the main-diagonal split of AbstractTetrahedralMesh::ConstructCuboid
_HEX_TETS = np.array([[0, 1, 5, 7], [0, 1, 3, 7], [0, 2, 3, 7], [0, 2, 6, 7], [0, 4, 6, 7], [0, 4, 5, 7]])


def ellipsoid_shell(h, a_in=2.5, b_in=1.0, a_out=3.5, b_out=2.0, theta_min=0.1, fibre_endo=-60.0,
                    fibre_epi=60.0):
    """(coords [n, 3], elems [ne, 4], fibres [ne, 3] unit vectors, apex [n] bool:
    the nodes on the apical rim theta = theta_min)."""
    thick = max(a_out - a_in, b_out - b_in)
    ns = max(2, int(round(thick / h)) + 1)
    mid_a, mid_b = 0.5 * (a_in + a_out), 0.5 * (b_in + b_out)
    arc = 0.5 * (mid_a + mid_b) * (np.pi / 2 - theta_min)
    nt = max(2, int(round(arc / h)) + 1)
    nphi = max(8, int(round(2 * np.pi * mid_b / h)))
    s = np.linspace(0.0, 1.0, ns)
    th = np.linspace(theta_min, np.pi / 2, nt)
    ph = np.arange(nphi) * (2 * np.pi / nphi)
    S, T, F = np.meshgrid(s, th, ph, indexing='ij')             # [ns, nt, nphi]
    a = a_in + S * (a_out - a_in)
    b = b_in + S * (b_out - b_in)
    coords = np.stack([b * np.sin(T) * np.cos(F), b * np.sin(T) * np.sin(F), -a * np.cos(T)], axis=-1).reshape(-1, 3)
    idx = np.arange(ns * nt * nphi).reshape(ns, nt, nphi)
    i0, j0, k0 = np.meshgrid(np.arange(ns - 1), np.arange(nt - 1), np.arange(nphi), indexing='ij')
    i0, j0, k0 = i0.ravel(), j0.ravel(), k0.ravel()
    k1 = (k0 + 1) % nphi
    corners = np.stack([idx[i0 + di, j0 + dj, kk]
                        for dk, kk in ((0, k0), (1, k1)) for dj in (0, 1) for di in (0, 1)], axis=1)
    # corner c has bits (s, theta, phi) = (c & 1, (c >> 1) & 1, c >> 2)
    elems = corners[:, _HEX_TETS].reshape(-1, 4).astype(np.int32)
    # fibre of each element at its centroid's parameters
    sc = np.repeat(s[i0] + 0.5 * (s[1] - s[0]), 6)
    tc = np.repeat(th[j0] + 0.5 * (th[1] - th[0]), 6)
    pc = np.repeat(ph[k0] + 0.5 * (2 * np.pi / nphi), 6)
    ac = a_in + sc * (a_out - a_in)
    bc = b_in + sc * (b_out - b_in)
    e_phi = np.stack([-np.sin(pc), np.cos(pc), np.zeros_like(pc)], axis=1)
    d_theta = np.stack([bc * np.cos(tc) * np.cos(pc), bc * np.cos(tc) * np.sin(pc), ac * np.sin(tc)], axis=1)
    e_theta = d_theta / np.linalg.norm(d_theta, axis=1, keepdims=True)
    alpha = np.radians(fibre_endo + sc * (fibre_epi - fibre_endo))
    fibres = np.cos(alpha)[:, None] * e_phi + np.sin(alpha)[:, None] * e_theta
    apex = np.zeros(coords.shape[0], dtype=bool)
    apex[idx[:, 0, :].ravel()] = True
    return coords, elems, fibres, apex


def disjoint_union(meshes):
    """One mesh holding several (coords, elems, fibres, apex) with no shared
    nodes: block-diagonal operators, so one context steps all of them."""
    coords, elems, fibres, apex, offsets = [], [], [], [], [0]
    for c, e, f, a in meshes:
        elems.append(e + offsets[-1])
        coords.append(c)
        fibres.append(f)
        apex.append(a)
        offsets.append(offsets[-1] + c.shape[0])
    return (np.concatenate(coords), np.concatenate(elems).astype(np.int32), np.concatenate(fibres),
            np.concatenate(apex), np.array(offsets))
