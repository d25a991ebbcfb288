"""Host-side DSS tables (csrc/ca_dss_table.cpp, built without HIP into
libcadycore_host.so): the per-column gather rows and the element clusters
the LDS-staged kernel indexes with.  Every index a kernel uses is checked
against the geometry's shared-node groups."""
import ctypes as C
import os
import types

import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import params as P

PAR = P.Params.from_toml()
HOST_LIB = os.path.join(os.path.dirname(G.__file__), "libcadycore_host.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HOST_LIB):
        pytest.skip("libcadycore_host.so not built")
    L = C.CDLL(HOST_LIB)
    L.ca_host_dss_build.restype = C.c_void_p
    L.ca_host_dss_build.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_char_p, C.c_int32]
    L.ca_host_dss_build_t.restype = C.c_void_p
    L.ca_host_dss_build_t.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]
    L.ca_host_dss_nclusters.argtypes = [C.c_void_p]
    L.ca_host_dss_get.argtypes = [C.c_void_p] + [C.c_void_p] * 9
    L.ca_host_dss_free.argtypes = [C.c_void_p]
    return L


def build(L, g, transpose=False):
    ptr = np.ascontiguousarray(g.dss_ptr, np.int32)
    col = np.ascontiguousarray(g.dss_col, np.int32)
    w = np.ascontiguousarray(g.dss_w, np.float64)
    basis = np.ascontiguousarray(g.basis, np.float64)
    err = C.create_string_buffer(256)
    h = L.ca_host_dss_build_t(g.nelem, g.ncol, len(ptr) - 1, ptr.ctypes.data, col.ctypes.data, w.ctypes.data,
                              basis.ctypes.data, int(transpose), err, 256)
    assert h, err.value
    ncl = L.ca_host_dss_nclusters(h)
    emax = L.ca_host_dss_emax()
    nc = g.ncol
    out = dict(tc=np.zeros(4 * nc, np.int32), tw=np.zeros(4 * nc, np.float32), tM=np.zeros(16 * nc, np.float32),
               list=np.zeros(ncl * emax, np.int32), nE=np.zeros(ncl, np.int32), ccol=np.zeros(ncl * 256, np.int32),
               slot=np.zeros(4 * ncl * 256, np.uint16), cw=np.zeros(4 * ncl * 256, np.float32),
               cM=np.zeros(16 * ncl * 256, np.float32))
    L.ca_host_dss_get(h, *[out[k].ctypes.data for k in ("tc", "tw", "tM", "list", "nE", "ccol", "slot", "cw", "cM")])
    L.ca_host_dss_free(h)
    out["ncl"], out["emax"] = ncl, emax
    return out


@pytest.mark.parametrize("h_elem", [2, 3, 8, 30])
def test_gather_rows_and_clusters_match_the_groups(lib, h_elem):
    g = G.cubed_sphere(h_elem, G.uniform_faces(4, 30000.0), PAR.planet_radius, PAR.Omega)
    T = build(lib, g)
    nc = g.ncol
    tc, tw = T["tc"].reshape(nc, 4), T["tw"].reshape(nc, 4)
    # gather rows: the group's members in group order, zero-weight padding = own column
    grp = -np.ones(nc, np.int64)
    for n in range(len(g.dss_ptr) - 1):
        grp[g.dss_col[g.dss_ptr[n]:g.dss_ptr[n + 1]]] = n
    for c in range(0, nc, 7):
        n = grp[c]
        if n < 0:
            assert tc[c, 0] == c and tw[c, 0] == 1.0 and np.all(tw[c, 1:] == 0)
            continue
        mem = g.dss_col[g.dss_ptr[n]:g.dss_ptr[n + 1]]
        m = len(mem)
        np.testing.assert_array_equal(tc[c, :m], mem)
        np.testing.assert_allclose(tw[c, :m], g.dss_w[g.dss_ptr[n]:g.dss_ptr[n + 1]], rtol=1e-7)
        assert np.all(tw[c, m:] == 0) and np.all(tc[c, m:] == c)
    assert np.allclose(tw.sum(axis=1), 1.0, atol=1e-6)
    # clusters: a partition of the columns; slots point at the same copies
    ncl, emax = T["ncl"], T["emax"]
    lst, nE = T["list"].reshape(ncl, emax), T["nE"]
    ccol, slot = T["ccol"].reshape(ncl, 256), T["slot"].reshape(ncl, 256, 4).astype(np.int64)
    owned = ccol[ccol >= 0]
    assert len(owned) == nc and len(np.unique(owned)) == nc
    assert np.all(nE <= emax) and np.all(nE >= 1)
    for cl in range(ncl):
        L = lst[cl, :nE[cl]]
        assert len(np.unique(L)) == len(L) and np.all((L >= 0) & (L < g.nelem))
        cols = ccol[cl]
        act = cols >= 0
        # own elements come first in the list, in thread order
        own_e = cols[act][::16] // 16
        np.testing.assert_array_equal(L[:len(own_e)], own_e)
        s = slot[cl][act]
        assert np.all(s < 16 * nE[cl])
        np.testing.assert_array_equal(16 * L[s // 16] + s % 16, tc[cols[act]])
        np.testing.assert_array_equal(T["cw"].reshape(ncl, 256, 4)[cl][act], tw[cols[act]])
    print(f"h{h_elem}: {ncl} clusters, staged elements per own element {nE.sum() / g.nelem:.2f}")


@pytest.mark.parametrize("h_elem", [8, 30])
def test_clusters_are_panel_blocks(lib, h_elem):
    """On panels numbered row by row the clusters are the 4 x 4 element blocks
    of each panel (narrower in the last block row / column), each staging at
    most its 6 x 6 neighbourhood."""
    g = G.cubed_sphere(h_elem, G.uniform_faces(4, 30000.0), PAR.planet_radius, PAR.Omega)
    T = build(lib, g)
    ccol, nE = T["ccol"].reshape(T["ncl"], 256), T["nE"]
    got = sorted(tuple(sorted(set((c[c >= 0] // 16).tolist()))) for c in ccol)
    want = []
    h = h_elem
    for p in range(6):
        for by in range(0, h, 4):
            for bx in range(0, h, 4):
                want.append(tuple(sorted(p * h * h + y * h + x for y in range(by, min(by + 4, h))
                                         for x in range(bx, min(bx + 4, h)))))
    assert got == sorted(want)
    assert nE.max() <= 36


def test_permuted_numbering_falls_back_to_grown_clusters(lib):
    """A numbering that is not panel by panel, row by row (elements shuffled)
    still gives valid clusters: a partition of the columns whose slots point
    at the same copies as the gather rows."""
    g0 = G.cubed_sphere(8, G.uniform_faces(4, 30000.0), PAR.planet_radius, PAR.Omega)
    rng = np.random.default_rng(8)
    perm = rng.permutation(g0.nelem)  # element e becomes perm[e]
    newcol = (16 * perm[np.arange(g0.ncol) // 16] + np.arange(g0.ncol) % 16).astype(np.int64)
    basis = np.empty_like(g0.basis)
    basis[:, newcol] = g0.basis
    g = types.SimpleNamespace(nelem=g0.nelem, ncol=g0.ncol, dss_ptr=g0.dss_ptr, dss_col=newcol[g0.dss_col],
                              dss_w=g0.dss_w, basis=basis)
    T = build(lib, g)
    nc, ncl, emax = g.ncol, T["ncl"], T["emax"]
    tc = T["tc"].reshape(nc, 4)
    ccol, slot = T["ccol"].reshape(ncl, 256), T["slot"].reshape(ncl, 256, 4).astype(np.int64)
    lst, nE = T["list"].reshape(ncl, emax), T["nE"]
    owned = ccol[ccol >= 0]
    assert len(owned) == nc and len(np.unique(owned)) == nc
    assert np.all(nE <= emax)
    for cl in range(ncl):
        cols = ccol[cl]
        act = cols >= 0
        s = slot[cl][act]
        np.testing.assert_array_equal(16 * lst[cl, :nE[cl]][s // 16] + s % 16, tc[cols[act]])
    # not the panel blocks of the original numbering
    assert ncl != 6 * 4


def test_independent_columns_have_no_clusters(lib):
    g = G.columns(5, G.uniform_faces(10, 30000.0))
    T = build(lib, g)
    assert T["ncl"] == 0
    tw = T["tw"].reshape(g.ncol, 4)
    assert np.all(tw[:, 0] == 1.0) and np.all(tw[:, 1:] == 0)


def test_bad_groups_are_rejected(lib):
    g = G.cubed_sphere(2, G.uniform_faces(4, 30000.0), PAR.planet_radius, PAR.Omega)
    ptr = np.ascontiguousarray(g.dss_ptr, np.int32)
    col = np.ascontiguousarray(g.dss_col, np.int32).copy()
    col[1] = col[0]  # a column in two slots
    w = np.ascontiguousarray(g.dss_w, np.float64)
    basis = np.ascontiguousarray(g.basis, np.float64)
    err = C.create_string_buffer(256)
    h = lib.ca_host_dss_build(g.nelem, g.ncol, len(ptr) - 1, ptr.ctypes.data, col.ctypes.data, w.ctypes.data,
                              basis.ctypes.data, err, 256)
    assert not h and b"partition" in err.value


def _apply(T, nc, x, v):
    """The gather DSS the kernels perform, from the tables: scalar x [nc],
    covariant vector v [nc, 2]."""
    tc, tw, tM = T["tc"].reshape(nc, 4), T["tw"].reshape(nc, 4).astype(np.float64), T["tM"].reshape(nc, 4, 2, 2)
    y = np.einsum("ct,ct->c", tw, x[tc])
    u = np.einsum("ctij,ctj->ci", tM.astype(np.float64), v[tc])
    return y, u


@pytest.mark.parametrize("h_elem", [2, 3])
def test_transposed_tables_are_the_adjoint_dss(lib, h_elem):
    """<DSS x, y> = <x, DSS^T y> for scalars and covariant vectors, and the
    transposed tables share the forward tables' slots and clusters."""
    g = G.cubed_sphere(h_elem, G.uniform_faces(4, 30000.0), PAR.planet_radius, PAR.Omega)
    F, B = build(lib, g), build(lib, g, transpose=True)
    nc = g.ncol
    for k in ("tc", "list", "nE", "ccol", "slot"):
        np.testing.assert_array_equal(F[k], B[k])
    rng = np.random.default_rng(h_elem)
    x, y = rng.standard_normal(nc), rng.standard_normal(nc)
    v, w = rng.standard_normal((nc, 2)), rng.standard_normal((nc, 2))
    Fx, Fv = _apply(F, nc, x, v)
    By, Bw = _apply(B, nc, y, w)
    assert abs(Fx @ y - x @ By) <= 1e-6 * np.sqrt((Fx @ Fx) * (y @ y))
    assert abs(np.sum(Fv * w) - np.sum(v * Bw)) <= 1e-6 * np.sqrt(np.sum(Fv * Fv) * np.sum(w * w))
    # and the forward DSS is a projection: applying it twice changes nothing (float32 tables)
    Fx2, Fv2 = _apply(F, nc, Fx, Fv)
    np.testing.assert_allclose(Fx2, Fx, atol=1e-5)
    np.testing.assert_allclose(Fv2, Fv, atol=1e-5)
    # cluster copies of the transposed weights and transforms
    ncl = F["ncl"]
    ccol = F["ccol"].reshape(ncl, 256)
    act = ccol >= 0
    np.testing.assert_array_equal(B["cw"].reshape(ncl, 256, 4)[act], B["tw"].reshape(nc, 4)[ccol[act]])
    np.testing.assert_array_equal(B["cM"].reshape(ncl, 256, 16)[act], B["tM"].reshape(nc, 16)[ccol[act]])
