"""The oracle's gradient limiters (src/BasicSolver.cpp:88-288) pinned by the
properties that define them (CPU only): Barth-Jespersen (limiter 1) keeps
every vertex reconstruction of a cell within the extrema of the cell and its
face neighbours, MLP (2) within the extrema of the cells around that vertex,
WENO (3) returns a convex combination of the cell's and its neighbours'
gradients; all three leave the exact gradient of a linear field unchanged
in interior cells, and Barth / MLP only scale a gradient by a factor in
[0, 1]."""
import numpy as np
import pytest

import sysu_cfd_heat_amd as P

TINY = 1e-10  # |reconstruction change| below which Barth_fun returns 1


def _solver(oracle, limiter, tets, dims=(5, 4, 4)):
    raw = P.box_mesh(*dims, 1.0, 1.0, 1.0, tets=tets, jitter=0.15 if tets else 0.0, seed=4,
                     side_rid=(1,) * 6)
    p = P.default_params()
    P.add_region(p, 1, 1, 1, 0.0)
    P.add_region(p, 5, 5, 0, 0.0, (0, 0, 0, 1, 0, 0.01))
    p.limiter = limiter
    return oracle.NSSolver(raw, p)


def _gradients(ns, u, bu, limiter):
    ns.set("Un", u)
    ns.set("BU", bu)
    ns.call("GradientU")
    return ns.get("dUdX")


def _raw_gradient(oracle, u, bu, tets):
    ns = _solver(oracle, 0, tets)
    g = _gradients(ns, u, bu, 0)
    ns.close()
    return g


@pytest.mark.parametrize("limiter", [1, 2, 3])
def test_linear_field_gradient_is_unchanged(oracle, limiter):
    # uniform hexahedra: Green-Gauss is exact for a linear field, the vertex
    # reconstructions stay within the neighbours' range and every gradient
    # is the same, so no limiter may change an interior cell's gradient
    ns = _solver(oracle, limiter, False, dims=(7, 6, 6))
    x, fx, bf = ns.mesh_get("cell_x"), ns.mesh_get("face_x"), ns.mesh_get("bnd_face")
    a = np.array([0.7, -1.1, 0.4])
    g = _gradients(ns, x @ a + 0.2, fx[bf] @ a + 0.2, limiter)
    interior = np.ones(len(x), bool)
    interior[ns.mesh_get("face_cell1")[bf]] = False
    if limiter == 3:  # WENO averages over face neighbours: they must be interior too
        cc = ns.mesh_get("cell_cell")
        interior = np.array([interior[i] and all(interior[j] for j in cc[i] if j >= 0)
                             for i in range(len(x))])
    assert interior.sum() >= 2
    np.testing.assert_allclose(g[interior], np.broadcast_to(a, g[interior].shape), rtol=0,
                               atol=1e-12)


@pytest.mark.parametrize("limiter", [1, 2])
@pytest.mark.parametrize("tets", [False, True])
def test_barth_and_mlp_bound_every_vertex_reconstruction(oracle, limiter, tets):
    ns = _solver(oracle, limiter, tets)
    n = ns.counts()[1]
    rng = np.random.default_rng(9 + limiter)
    u = rng.standard_normal(n)
    bu = rng.standard_normal(ns.counts()[3])
    g0 = _raw_gradient(oracle, u, bu, tets)
    g = _gradients(ns, u, bu, limiter)
    x, vert, cv = ns.mesh_get("cell_x"), ns.mesh_get("vert"), ns.mesh_get("cell_vert")
    # a scalar factor in [0, 1] per cell
    f = np.einsum("ij,ij->i", g, g0) / np.maximum(np.einsum("ij,ij->i", g0, g0), 1e-300)
    np.testing.assert_allclose(g, f[:, None] * g0, rtol=1e-12, atol=1e-14)
    assert np.all(f >= -1e-14) and np.all(f <= 1 + 1e-14)
    assert (f < 0.999).sum() > n // 10  # the limiter acts on a random field
    if limiter == 1:  # extrema over the cell and its face neighbours
        cc = ns.mesh_get("cell_cell")
        lo = np.array([min([u[i]] + [u[j] for j in cc[i] if j >= 0]) for i in range(n)])
        hi = np.array([max([u[i]] + [u[j] for j in cc[i] if j >= 0]) for i in range(n)])
        lo_v = np.repeat(lo[:, None], 8, 1)
        hi_v = np.repeat(hi[:, None], 8, 1)
    else:  # extrema over the cells around each vertex
        vmin = np.full(len(vert), np.inf)
        vmax = np.full(len(vert), -np.inf)
        np.minimum.at(vmin, cv.ravel(), np.repeat(u, 8))
        np.maximum.at(vmax, cv.ravel(), np.repeat(u, 8))
        lo_v, hi_v = vmin[cv], vmax[cv]
    rec = u[:, None] + np.einsum("ivk,ik->iv", vert[cv] - x[:, None, :], g)
    slack = TINY + 1e-12 * np.abs(u).max()
    assert np.all(rec <= hi_v + slack) and np.all(rec >= lo_v - slack)


@pytest.mark.parametrize("tets", [False, True])
def test_weno_is_a_convex_combination_of_neighbour_gradients(oracle, tets):
    ns = _solver(oracle, 3, tets)
    n = ns.counts()[1]
    rng = np.random.default_rng(21)
    u = rng.standard_normal(n)
    bu = rng.standard_normal(ns.counts()[3])
    g0 = _raw_gradient(oracle, u, bu, tets)
    g = _gradients(ns, u, bu, 3)
    cc = ns.mesh_get("cell_cell")
    for i in range(n):
        pool = np.array([g0[i]] + [g0[j] for j in cc[i] if j >= 0])
        assert np.all(g[i] >= pool.min(0) - 1e-12) and np.all(g[i] <= pool.max(0) + 1e-12)
        w = 1.0 / (np.einsum("ij,ij->i", pool, pool) + 1e-16)
        np.testing.assert_allclose(g[i], (w[:, None] * pool).sum(0) / w.sum(), rtol=1e-10,
                                   atol=1e-14)
