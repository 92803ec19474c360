"""Writes the golden fixtures of tests/golden/ (run from the repository root:
python tests/golden/make_golden.py). Inputs come from the deterministic case
generators; expected outputs from the oracle (the CPU restatement, FP64) and,
for the linear systems, from dense LAPACK solves. The reference itself holds
no fixtures for this path and cannot be built here (LASPack is absent), so
these pin regressions of both the product and the oracle; the independent
pins are the closed-form and invariant tests."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import sysu_cfd_heat_amd as P  # noqa: E402
from sysu_cfd_heat_amd import cases  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

MESH_INT = ("face_cell1", "face_cell2", "face_bnd", "cell_nface", "cell_face", "cell_cell",
            "bnd_face")
MESH_DBL = ("face_area", "face_lambda", "face_rlencos", "face_n", "face_x", "face_xpac",
            "face_xnac", "cell_x", "cell_vol", "bnd_distance")


def mesh_fixture():
    raw = P.box_mesh(3, 2, 2, tets=True, jitter=0.1, seed=3)
    m = O.Mesh(raw)
    d = {"in_vert": raw.vert, "in_cell_vert": raw.cell_vert, "in_cell_rid": raw.cell_rid,
         "in_bnd_vert": raw.bnd_vert, "in_bnd_rid": raw.bnd_rid,
         "counts": np.array(m.counts, np.int32)}
    for nm in MESH_INT + MESH_DBL:
        d[nm] = m.get(nm)
    np.savez_compressed(os.path.join(OUT, "mesh_tet_3x2x2.npz"), **d)


def simple_fixture():
    raw, p = cases.cavity(n=4, tets=True, jitter=0.1, seed=5, max_step=5)
    ref = O.NSSolver(raw, p, mc=1)
    d = {}
    for it in range(3):
        ref.call("Iterate")
    for nm in ("Un", "Vn", "Wn", "Pn", "RUFace"):
        d[nm] = ref.get(nm)
    d["Residual"] = ref.get("Residual")
    np.savez_compressed(os.path.join(OUT, "cavity_tet4_iter3.npz"), **d)


def las_fixture():
    from test_oracle_las import convection_diffusion, laplacian_3d
    A = laplacian_3d(4, 0.05)
    N = convection_diffusion(4, 2.0)
    rng = np.random.default_rng(11)
    bs, bn = rng.standard_normal(A.shape[0]), rng.standard_normal(N.shape[0])
    np.savez_compressed(os.path.join(OUT, "las_small.npz"), A_spd=A, b_spd=bs,
                        x_spd=np.linalg.solve(A, bs), A_nonsym=N, b_nonsym=bn,
                        x_nonsym=np.linalg.solve(N, bn))


if __name__ == "__main__":
    mesh_fixture()
    simple_fixture()
    las_fixture()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)), "bytes")
