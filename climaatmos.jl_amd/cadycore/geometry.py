"""Grid geometry of the dry-dycore step (host side, Float64).

The native library takes geometry as an input, the way ClimaAtmos takes
ClimaCore's ``LocalGeometry``: a Julia caller passes ClimaCore's own metric
terms, a standalone caller builds them here.  Everything is stored *factored*:

* per stored column (``ncol = 16 * nelem``): the horizontal metric at the
  reference radius R -- Jacobian ``J_h``, inverse metric ``g^{ij}``, metric
  ``g_{ij}``, Coriolis factors, and the Cartesian covariant/contravariant basis
  vectors used by the vector DSS;
* per level: face heights ``z_face``; centres, vertical Jacobians and the
  deep-atmosphere scale ``s(z) = (R + z) / R`` are derived from them.

The full 3-D metric is the product: ``J = J_h s^2 Jv``, ``g^{ij}_h(z) =
g^{ij}_h / s^2``, ``g^{33} = 1 / Jv^2`` (no topography).

Conventions restated from the reference grid constructors
(``src/simulation/grids.jl:42-129``): equiangular cubed sphere with
``h_elem`` elements per panel edge, Gauss-Lobatto-Legendre nodes with
``nh_poly + 1 = 4`` points, no bubble correction, deep atmosphere, and
``HyperbolicTangentStretching(dz_bottom)`` in the vertical
(``grids.jl:356-357``).  The vertical coordinate of the finite-difference
operators is the level index: the centre Jacobian is the layer thickness and
the face Jacobian is the centre-to-centre distance (half layers extended to a
full layer at the two boundary faces).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

NQ = 4


def gll(nq: int = NQ):
    """GLL nodes, weights and derivative matrix ``D[i, k] = l_k'(xi_i)``."""
    if nq != 4:
        raise ValueError("only nq = 4 is supported")
    s5 = 1.0 / math.sqrt(5.0)
    x = np.array([-1.0, -s5, s5, 1.0])
    w = np.array([1.0 / 6.0, 5.0 / 6.0, 5.0 / 6.0, 1.0 / 6.0])
    n = nq - 1

    def legendre(n, t):
        p0, p1 = 1.0, t
        if n == 0:
            return p0
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * t * p1 - (k - 1) * p0) / k
        return p1

    L = np.array([legendre(n, t) for t in x])
    D = np.zeros((nq, nq))
    for i in range(nq):
        for k in range(nq):
            if i != k:
                D[i, k] = (L[i] / L[k]) / (x[i] - x[k])
    D[0, 0] = -n * (n + 1) / 4.0
    D[n, n] = n * (n + 1) / 4.0
    return x, w, D


def tanh_stretched_faces(z_elem: int, z_max: float, dz_bottom: float) -> np.ndarray:
    """Face heights of ``HyperbolicTangentStretching(dz_bottom)``.

    ``z_k = z_max * eta(k / N)``, ``eta(zeta) = 1 - tanh(g (1 - zeta)) / tanh(g)``
    with ``g`` chosen so that the bottom layer is ``dz_bottom`` thick (solved
    by bisection to machine precision).
    """
    if z_elem < 2:
        raise ValueError("z_elem must be > 1")
    target = dz_bottom / z_max
    if not (0.0 < target < 1.0 / z_elem):
        raise ValueError("dz_bottom must be smaller than the mean layer thickness")
    zeta1 = 1.0 / z_elem

    def eta(zeta, g):
        return 1.0 - math.tanh(g * (1.0 - zeta)) / math.tanh(g)

    lo, hi = 1e-8, 1.0
    while eta(zeta1, hi) > target:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if eta(zeta1, mid) > target:
            lo = mid
        else:
            hi = mid
    g = 0.5 * (lo + hi)
    k = np.arange(z_elem + 1) / z_elem
    z = z_max * (1.0 - np.tanh(g * (1.0 - k)) / math.tanh(g))
    z[0] = 0.0
    z[-1] = z_max
    return z


def uniform_faces(z_elem: int, z_max: float) -> np.ndarray:
    return np.linspace(0.0, z_max, z_elem + 1)


def vertical_metrics(z_face: np.ndarray):
    """Centre heights and the index-coordinate vertical Jacobians."""
    zf = np.asarray(z_face, dtype=np.float64)
    zc = 0.5 * (zf[1:] + zf[:-1])
    Jv_c = zf[1:] - zf[:-1]
    Jv_f = np.empty_like(zf)
    Jv_f[0] = zf[1] - zf[0]
    Jv_f[-1] = zf[-1] - zf[-2]
    Jv_f[1:-1] = zc[1:] - zc[:-1]
    return zc, Jv_c, Jv_f


# Cube panel frames (normal, e_alpha, e_beta), all right-handed with
# e_alpha x e_beta = normal so that every element has J > 0.
_PANELS = [
    ((1, 0, 0), (0, 1, 0), (0, 0, 1)),
    ((0, 1, 0), (-1, 0, 0), (0, 0, 1)),
    ((-1, 0, 0), (0, -1, 0), (0, 0, 1)),
    ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
    ((0, 0, 1), (0, 1, 0), (-1, 0, 0)),
    ((0, 0, -1), (0, 1, 0), (1, 0, 0)),
]


@dataclass
class Grid:
    """Factored geometry of a sphere or of a set of independent columns."""

    kind: str  # "sphere" or "columns"
    nelem: int
    ncol: int
    nlev: int
    radius: float  # deep-atmosphere radius; 0 disables the s(z) scaling
    z_face: np.ndarray  # (nlev + 1,)
    J_h: np.ndarray  # (ncol,)
    gup: np.ndarray  # (3, ncol): g^11, g^12, g^22 at radius R
    gdn: np.ndarray  # (3, ncol): g_11, g_12, g_22 at radius R
    f3_num: np.ndarray  # (ncol,): vertical Coriolis 2 Omega sin(lat) (or f)
    f12: np.ndarray  # (2, ncol): contravariant horizontal Coriolis at radius R
    basis: np.ndarray  # (12, ncol): a^1, a^2, a_1, a_2 (Cartesian, radius R)
    lat: np.ndarray  # (ncol,) radians
    lon: np.ndarray  # (ncol,) radians
    node_id: np.ndarray  # (ncol,) unique horizontal node of each column
    n_nodes: int
    dss_ptr: np.ndarray  # CSR over shared nodes (int32)
    dss_col: np.ndarray
    dss_w: np.ndarray
    h: float  # mean nodal spacing [m]
    h_elem: int = 0

    @property
    def zc(self):
        return vertical_metrics(self.z_face)[0]

    def vertical(self):
        return vertical_metrics(self.z_face)

    def scale(self, z):
        if self.radius > 0.0:
            return (self.radius + np.asarray(z)) / self.radius
        return np.ones_like(np.asarray(z, dtype=np.float64))


def _dss_groups(node_id: np.ndarray, n_nodes: int, J_h: np.ndarray):
    order = np.argsort(node_id, kind="stable")
    counts = np.bincount(node_id, minlength=n_nodes)
    starts = np.concatenate([[0], np.cumsum(counts)])
    ptr = [0]
    cols = []
    ws = []
    for n in range(n_nodes):
        members = order[starts[n]:starts[n + 1]]
        if len(members) < 2:
            continue
        jw = J_h[members]
        cols.extend(members.tolist())
        ws.extend((jw / jw.sum()).tolist())
        ptr.append(len(cols))
    return (np.asarray(ptr, dtype=np.int32), np.asarray(cols, dtype=np.int32),
            np.asarray(ws, dtype=np.float64))


def cubed_sphere(h_elem: int, z_face: np.ndarray, radius: float, Omega: float,
                 deep: bool = True) -> Grid:
    """Equiangular cubed sphere with ``h_elem`` elements per panel edge.

    Element order: panel, then ``ey``, then ``ex``; column index
    ``col = 16 * e + 4 * j + i`` with ``i`` along the first reference
    coordinate.
    """
    xq, wq, D = gll()
    nelem = 6 * h_elem * h_elem
    ncol = 16 * nelem
    dang = (math.pi / 2.0) / h_elem
    cart = np.empty((3, ncol))
    a1 = np.empty((3, ncol))
    a2 = np.empty((3, ncol))
    e = 0
    for p, (nrm, ea, eb) in enumerate(_PANELS):
        nrm = np.array(nrm, float)
        ea = np.array(ea, float)
        eb = np.array(eb, float)
        for ey in range(h_elem):
            for ex in range(h_elem):
                for j in range(NQ):
                    for i in range(NQ):
                        col = 16 * e + 4 * j + i
                        al = -math.pi / 4 + (ex + 0.5 * (xq[i] + 1.0)) * dang
                        be = -math.pi / 4 + (ey + 0.5 * (xq[j] + 1.0)) * dang
                        ta, tb = math.tan(al), math.tan(be)
                        c = nrm + ta * ea + tb * eb
                        nc = math.sqrt(c @ c)
                        chat = c / nc
                        P = (np.eye(3) - np.outer(chat, chat)) / nc
                        d1 = P @ ea * (1.0 + ta * ta) * (0.5 * dang)
                        d2 = P @ eb * (1.0 + tb * tb) * (0.5 * dang)
                        cart[:, col] = chat
                        a1[:, col] = radius * d1
                        a2[:, col] = radius * d2
                e += 1
    g11 = np.einsum("ic,ic->c", a1, a1)
    g12 = np.einsum("ic,ic->c", a1, a2)
    g22 = np.einsum("ic,ic->c", a2, a2)
    det = g11 * g22 - g12 * g12
    J_h = np.einsum("ic,ic->c", np.cross(a1.T, a2.T).T, cart)
    gu11, gu12, gu22 = g22 / det, -g12 / det, g11 / det
    A1 = gu11 * a1 + gu12 * a2  # contravariant basis a^1
    A2 = gu12 * a1 + gu22 * a2
    zhat = np.array([0.0, 0.0, 1.0])
    lat = np.arcsin(np.clip(cart[2], -1.0, 1.0))
    lon = np.arctan2(cart[1], cart[0])
    f3_num = 2.0 * Omega * cart[2]
    if deep:
        f12 = 2.0 * Omega * np.stack([A1.T @ zhat, A2.T @ zhat])
    else:
        f12 = np.zeros((2, ncol))
    key = np.round(cart * 1e9).astype(np.int64)
    _, node_id = np.unique(key.T, axis=0, return_inverse=True)
    node_id = node_id.reshape(-1).astype(np.int64)
    # renumber unique nodes by first occurrence
    first = {}
    remap = np.empty(node_id.max() + 1, dtype=np.int64)
    nxt = 0
    for c in range(ncol):
        n = node_id[c]
        if n not in first:
            first[n] = nxt
            remap[n] = nxt
            nxt += 1
    node_id = remap[node_id]
    n_nodes = nxt
    dss_ptr, dss_col, dss_w = _dss_groups(node_id, n_nodes, J_h)
    from .params import node_horizontal_length_scale

    basis = np.concatenate([A1, A2, a1, a2], axis=0)
    nlev = len(z_face) - 1
    return Grid(
        kind="sphere", nelem=nelem, ncol=ncol, nlev=nlev,
        radius=radius if deep else 0.0, z_face=np.asarray(z_face, float),
        J_h=J_h, gup=np.stack([gu11, gu12, gu22]), gdn=np.stack([g11, g12, g22]),
        f3_num=f3_num, f12=f12, basis=basis, lat=lat, lon=lon,
        node_id=node_id, n_nodes=n_nodes, dss_ptr=dss_ptr, dss_col=dss_col,
        dss_w=dss_w, h=node_horizontal_length_scale(radius, h_elem), h_elem=h_elem,
    )


def columns(ncol: int, z_face: np.ndarray, f_plane: float = 0.0) -> Grid:
    """``ncol`` independent single columns (``ColumnGrid``): unit horizontal
    metric, f-plane Coriolis ``f`` and no deep-atmosphere scaling."""
    one = np.ones(ncol)
    zero = np.zeros(ncol)
    nlev = len(z_face) - 1
    return Grid(
        kind="columns", nelem=0, ncol=ncol, nlev=nlev, radius=0.0,
        z_face=np.asarray(z_face, float), J_h=one.copy(),
        gup=np.stack([one, zero, one]), gdn=np.stack([one, zero, one]),
        f3_num=np.full(ncol, float(f_plane)), f12=np.zeros((2, ncol)),
        basis=np.zeros((12, ncol)), lat=zero.copy(), lon=zero.copy(),
        node_id=np.arange(ncol), n_nodes=ncol,
        dss_ptr=np.zeros(1, np.int32), dss_col=np.zeros(0, np.int32),
        dss_w=np.zeros(0), h=float("inf"),
    )
