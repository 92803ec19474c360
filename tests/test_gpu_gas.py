"""Perfect gas (DensityModel 1, rho = p / (Rcpcv T)) on the device against
the oracle: the non-symmetric pressure matrix, the density correction of the
face fluxes and the viscous-work energy source bit for bit on the same state
(no global reduction enters them: DensityModel 1 has no outlet mass
balance), SIMPLE iterations and time steps within the solvers' tolerance,
and the Mach column of the Tecplot output."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pair(steady=True, dt=0.05, scheme=1, turbulence=False, t0=1.0, bicgstab=False):
    import sysu_cfd_heat_amd as P
    from oracle import oracle as O
    from sysu_cfd_heat_amd import cases
    raw, p = cases.gas_channel(nx=10, ny=4, nz=4, tets=True, jitter=0.1, seed=2,
                               turbulence=turbulence, max_step=1000, t0=t0, t_wall=1.2 * t0)
    if bicgstab:
        p.pres = P.solve_cfg(P.CFD_BICGSTAB, P.CFD_PC_ILU, 1.0, 1e-8, 2000)
    if not steady:
        p.steady, p.dt, p.time_scheme, p.max_outer_step = 0, dt, scheme, 3
    return P.NSSolver(raw, p), O.NSSolver(raw, p, mc=1)


def _random_state(gpu, ref, seed):
    rng = np.random.default_rng(seed)
    n, nf = gpu.counts()[1], gpu.counts()[2]
    vals = {nm: rng.standard_normal(n) for nm in ("Un", "Vn", "Wn", "Pn")}
    vals["Tn"] = 1.0 + 0.5 * rng.random(n)
    vals["Rn"] = 0.8 + 0.4 * rng.random(n)
    vals["VisTur"] = 1e-3 * rng.random(n)
    vals["Apr"] = 0.5 + rng.random(n)
    for nm, v in vals.items():
        gpu.set(nm, v)
        ref.set(nm, v)
    f = 0.05 * rng.standard_normal(nf)
    gpu.set("RUFace", f)
    ref.set("RUFace", f)


def _same(a, b, what):
    assert np.array_equal(a, b), f"{what}: max diff {np.abs(a - b).max():.3e}"


@pytest.mark.parametrize("steady", [True, False])
def test_gas_pressure_matrix_bit_identical(steady):
    print(f"[gas] pressure matrix steady={steady}", flush=True)
    gpu, ref = _pair(steady)
    _random_state(gpu, ref, 1)
    for s in (gpu, ref):
        s.call("BuildVelocityMatrix")
        s.call("BuildPressureMatrix")
    _same(gpu.get("RUFace"), ref.get("RUFace"), "RUFace")
    _same(gpu.get("bp"), ref.get("bp"), "bp")
    ga, ra = gpu.matrix("Ap"), ref.matrix("Ap")
    assert np.array_equal(ga[0], ra[0]) and np.array_equal(ga[1], ra[1])
    _same(ga[2], ra[2], "Ap values")
    n = gpu.counts()[1]
    assert ga[0].sum() > 2 * n  # full rows: every neighbour in every row


def test_gas_flux_correction_and_energy_source_bit_identical():
    print("[gas] CorrectRUFace2 + EnergyViscSource", flush=True)
    gpu, ref = _pair()
    _random_state(gpu, ref, 2)
    dp = np.random.default_rng(3).standard_normal(gpu.counts()[1])
    for s in (gpu, ref):
        s.call("BuildVelocityMatrix")
        s.set("xsol", dp)
        s.call("SetBCDeltaP")
        s.call("GradientP")
        s.call("SetBCVelocity")
        s.call("CorrectRUFace2")
        for g in ("GradientU", "GradientV", "GradientW"):
            s.call(g)
        s.call("EnergyViscSource")
    _same(gpu.get("RUFace"), ref.get("RUFace"), "RUFace corrected")
    _same(gpu.get("ESource"), ref.get("ESource"), "ESource")


# Steady perfect-gas SIMPLE is fragile in the reference formulation: at
# t0 = 1 (Mach ~ 0.05) the nearly singular pressure system stalls GMRES, and
# where a solve converges the density runs away within a few iterations. The
# comparison runs two iterations at t0 = 0.05 with BiCGSTAB, where every solve
# converges and the density stays positive.
@pytest.mark.parametrize("turbulence", [False, True])
def test_gas_simple_iterations_match_oracle(turbulence):
    print(f"[gas] iterations turbulence={turbulence}", flush=True)
    gpu, ref = _pair(turbulence=turbulence, t0=0.05, bicgstab=True)
    for _ in range(2):
        gpu.call("Iterate")
        ref.call("Iterate")
    names = ["Un", "Vn", "Wn", "Pn", "Rn", "Tn"] + (["TE", "ED"] if turbulence else [])
    for nm in names:
        e = rel_l2(gpu.get(nm), ref.get(nm))
        assert e <= 1e-6, f"{nm}: rel-L2 {e:.3e}"
    it, rr = gpu.last_solve("p")
    assert rr < 1e-7, (it, rr)
    # the density moved with the pressure correction: not the initial field
    assert np.abs(gpu.get("Rn") - 1.0).max() > 1e-6


@pytest.mark.parametrize("scheme", [1, 2])
def test_gas_time_steps_match_oracle(scheme):
    print(f"[gas] transient scheme={scheme}", flush=True)
    gpu, ref = _pair(steady=False, scheme=scheme)
    for _ in range(2):
        for s in (gpu, ref):
            s.call("SaveTransientOldData")
            for _ in range(2):
                s.call("Iterate")
    for nm in ("Un", "Vn", "Wn", "Pn", "Rn", "Tn"):
        e = rel_l2(gpu.get(nm), ref.get(nm))
        assert e <= 1e-6, f"{nm}: rel-L2 {e:.3e}"
    it, rr = gpu.last_solve("p")  # GMRES (the configuration names CG)
    assert rr < 1e-7 and it < 500, (it, rr)


def test_gas_tecplot_writes_mach(tmp_path):
    print("[gas] tecplot Mach", flush=True)
    gpu, _ = _pair()
    for _ in range(2):
        gpu.call("Iterate")
    path = str(tmp_path / "gas.dat")
    gpu.write_tecplot(path)
    lines = open(path).read().split("\n")
    assert lines[0] == 'variables="x","y","z""p","u","v","w","ro","T","Mach","mu"'
    toks = " ".join(lines[1:]).split()
    nv, nc = int(toks[1][2:-1]), int(toks[2][2:-1])
    k = 5 + 3 * nv
    cellv = np.array(toks[k:k + 8 * nc], float).reshape(8, nc)
    u, v, w = gpu.get("Un"), gpu.get("Vn"), gpu.get("Wn")
    from sysu_cfd_heat_amd import cases
    _, p = cases.gas_channel(nx=10, ny=4, nz=4)
    mach = np.sqrt((u * u + v * v + w * w) / (1.4 * (gpu.get("Pn") + p.pressure_ref)
                                              / gpu.get("Rn")))
    np.testing.assert_allclose(cellv[6], mach, rtol=1e-5, atol=1e-12)
    assert 0 < mach.max() < 0.2  # low Mach: |u| ~ 1 against c ~ 20
