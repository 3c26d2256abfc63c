"""Parity of the HIP solver (through the C ABI) with the reference and the oracle.

All-FP64 path: the golden right-hand sides, face values and time-loop runs recorded from
the reference's unmodified 4-rank solver are reproduced bit for bit (tolerance 0), and so
are the oracle's results on further states, grids and wall boundaries."""
import numpy as np
import pytest

import oracle as orc
import wenofv
from conftest import golden

pytestmark = pytest.mark.gpu

RHS = golden("solver_rhs.npz")
RUNS = golden("solver_runs.npz")
CASE = {
    0: ((0.0, 1.0, 0.0, 1.0), 1.2, 1.0, 1.0, 1.0, 0.0),
    1: ((-1.0, 1.0, -1.0, 1.0), 1000.0, 1.0, 30.0, 30.0, 0.0),
    2: ((0.0, 1.0, 0.0, 1.0), 1000.0, 1.0, 70.0, 70.0, 0.0),
}


def rhs_tags():
    return sorted(k[:-4] for k in RHS.files if k.endswith("_rhs"))


@pytest.mark.parametrize("tag", rhs_tags())
def test_rhs_and_faces_match_reference_bitwise(tag):
    kind, s, r, f, b = tag.split("_")
    p = RHS[kind + "_params"]
    U = RHS[kind + "_U"]
    ny, nx = U.shape[1:]
    print(f"case {tag} {nx}x{ny}", flush=True)
    sol = wenofv.Solver(nx, ny, rho10=p[0], rho20=p[1], sound1=p[2], sound2=p[3], preb=p[4], scheme=int(s[1]),
                        recon=int(r[1]), flux=int(f[1]), bp=int(b[1]))
    sol.set_state(U)
    rhs, faces, theta = sol.assemble_rhs(want_faces=True)
    np.testing.assert_array_equal(rhs, RHS[tag + "_rhs"])
    if tag + "_faces" in RHS.files:
        np.testing.assert_array_equal(faces, RHS[tag + "_faces"])
    if int(b[1]):
        np.testing.assert_array_equal((theta < 1 - 1e-8).astype(float), RHS[tag + "_thflag"])


def run_names():
    return sorted(k[:-5] for k in RUNS.files if k.endswith("_opts"))


@pytest.mark.parametrize("name", run_names())
def test_time_loop_matches_reference_bitwise(name):
    tc, nx, ny, rk, cfl, scheme, bp, flux, recon, fixed, tend = RUNS[name + "_opts"]
    dom, rho10, rho20, c1, c2, preb = CASE[int(tc)]
    print(f"run {name}", flush=True)
    sol = wenofv.Solver(int(nx), int(ny), domain=dom, rho10=rho10, rho20=rho20, sound1=c1, sound2=c2, preb=preb,
                        scheme=int(scheme), recon=int(recon), flux=int(flux), bp=int(bp), rk=int(rk), cfl=cfl)
    ref_dts = RUNS[name + "_dts"]
    # device-resident loop (graph replay)
    sol.set_state(RUNS[name + "_init"])
    if fixed:
        n, t, dts = sol.advance(fixed_steps=int(tend), max_dts=len(ref_dts) + 4)
    else:
        n, t, dts = sol.advance(t_end=tend, max_dts=len(ref_dts) + 4)
    np.testing.assert_array_equal(dts, ref_dts)
    np.testing.assert_array_equal(sol.get_state(), RUNS[name + "_final"])
    # host-driven loop: calculateDeltaT + performTimeIntegration
    sol.set_state(RUNS[name + "_init"])
    now = 0.0
    for k in range(len(ref_dts)):
        dt = sol.compute_dt()
        if not fixed:
            dt = min(dt, tend - now)
        assert dt == ref_dts[k]
        sol.step(dt)
        now += dt
    np.testing.assert_array_equal(sol.get_state(), RUNS[name + "_final"])
    assert sol.flags() == 0


def random_state(rng, nx, ny, params, sharp):
    rho10, rho20, c1, c2, preb = params
    X, Y = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny)
    if sharp:
        z1 = np.where((X - 0.45) ** 2 + (Y - 0.55) ** 2 < 0.09, 1 - 1e-6, 1e-6)
        pr = preb + 1.0 + 0.5 * (X < 0.15)
        u = 0.3 * (X < 0.15) + 0.05 * rng.standard_normal((ny, nx))
        v = 0.05 * rng.standard_normal((ny, nx))
    else:
        z1 = 0.5 + 0.3 * np.sin(2 * np.pi * X) * np.cos(4 * np.pi * Y) + 0.03 * rng.standard_normal((ny, nx))
        pr = preb + 0.5 + 0.05 * rng.standard_normal((ny, nx))
        u = 0.4 + 0.1 * rng.standard_normal((ny, nx))
        v = -0.2 + 0.1 * rng.standard_normal((ny, nx))
    r1, r2 = rho10 + (pr - preb) / c1 ** 2, rho20 + (pr - preb) / c2 ** 2
    a, b = z1 * r1, (1 - z1) * r2
    return np.stack([a, b, (a + b) * u, (a + b) * v])


COMBOS = [(s, r, f, b) for s in (0, 1) for r in (0, 1, 2, 3) for f in (0, 1) for b in (0, 1)]


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("combo", COMBOS)
def test_rhs_matches_oracle_on_ragged_grids(combo, bc):
    s, r, f, b = combo
    rng = np.random.default_rng(hash((combo, bc)) % 2**32)
    nx, ny = (70, 37) if bc else (64, 45)   # strips of 58 columns: one full and one partial
    params = (1.0, 0.1, 10.0, 10.0, 0.0)
    sharp = bool(b) and r < 2
    U = random_state(rng, nx, ny, params, sharp)
    print(f"oracle rhs s{s} r{r} f{f} b{b} bc{bc} {nx}x{ny}", flush=True)
    kw = dict(scheme=s, recon=r, flux=f, bp=b, rho10=params[0], rho20=params[1], sound1=params[2],
              sound2=params[3], preb=params[4], domain=(0.0, 2.0, 0.0, 1.0))
    o = orc.Oracle(nx, ny, bc=bc, nranks=1, **kw)
    sol = wenofv.Solver(nx, ny, bc=bc, geometry_ranks=1, rows_per_strip=8, **kw)
    sol.set_state(U)
    rhs, faces, theta = sol.assemble_rhs(want_faces=True)
    r_o, f_o, t_o, err = o.assemble_rhs(U, want_faces=True)
    np.testing.assert_array_equal(rhs, r_o)
    np.testing.assert_array_equal(faces, f_o)
    np.testing.assert_array_equal(theta, t_o)


@pytest.mark.parametrize("rk", [0, 1, 2])
def test_walls_time_loop_matches_oracle(rk):
    rng = np.random.default_rng(7 + rk)
    nx, ny = 96, 40
    params = (1.0, 1e-3, 10.0, 10.0, 0.0)
    U = random_state(rng, nx, ny, params, True)
    kw = dict(scheme=1, recon=0, flux=1, bp=1, rk=rk, cfl=0.4, rho10=params[0], rho20=params[1],
              sound1=params[2], sound2=params[3], preb=params[4], domain=(0.0, 2.0, 0.0, 1.0))
    o = orc.Oracle(nx, ny, bc=1, nranks=1, **kw)
    sol = wenofv.Solver(nx, ny, bc=1, geometry_ranks=1, **kw)
    sol.set_state(U)
    n, t, dts = sol.advance(fixed_steps=12, max_dts=16)
    U_o, dts_o, t_o, err = o.run(U, fixed_steps=12)
    assert err == 0 and n == 12
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(sol.get_state(), U_o)
    # reflective walls conserve the mass of each phase
    m0 = U[:2].sum(axis=(1, 2))
    np.testing.assert_allclose(sol.get_state()[:2].sum(axis=(1, 2)), m0, rtol=1e-12)
