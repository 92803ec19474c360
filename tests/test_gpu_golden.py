"""The device path against the committed golden fixtures (no oracle run)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_simple_iterations_match_fixture():
    import sysu_cfd_heat_amd as P
    from sysu_cfd_heat_amd import cases
    g = np.load(os.path.join(GOLD, "cavity_tet4_iter3.npz"))
    raw, p = cases.cavity(n=4, tets=True, jitter=0.1, seed=5, max_step=5)
    s = P.NSSolver(raw, p)
    for _ in range(3):
        s.call("Iterate")
    for nm in ("Un", "Vn", "Wn", "Pn", "RUFace"):
        e = rel_l2(s.get(nm), g[nm])
        assert e <= 1e-6, (nm, e)
    s.close()


@pytest.mark.parametrize("kind,method,pc", [("spd", "CGIter", "ILUPrecond"),
                                            ("spd", "CGIter", "JacobiPrecond"),
                                            ("nonsym", "BiCGSTABIter", "ILUPrecond"),
                                            ("nonsym", "GMRESIter", "SSORPrecond")])
def test_laspack_abi_solves_fixture(kind, method, pc):
    from test_gpu_laspack import _get, _las, _matrix, _vector
    g = np.load(os.path.join(GOLD, "las_small.npz"))
    A, b, x = g["A_" + kind], g["b_" + kind], g["x_" + kind]
    L = _las()
    Q = _matrix(L, A, sym=(kind == "spd"))
    xv, bv = _vector(L, np.zeros(len(b))), _vector(L, b)
    L.SetRTCAccuracy(1e-12)
    getattr(L, method)(C.byref(Q), C.byref(xv), C.byref(bv), 2000,
                       C.cast(getattr(L, pc), C.c_void_p), 1.2)
    assert L.LASResult() == 0 and L.GetLastAccuracy() < 1e-12
    assert np.linalg.norm(_get(xv) - x) <= 1e-9 * np.linalg.norm(x)
    L.Q_Destr(C.byref(Q)); L.V_Destr(C.byref(xv)); L.V_Destr(C.byref(bv))
