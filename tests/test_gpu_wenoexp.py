"""The 7-point forms on the GPU: SCHEME 2 (WENO-EXP) with every reconstruction form and
RECONSTRUCTIONTYPE 4 (primitive3: THINC on z1, WENO-Z on p, u, v).

Against the reference's own right-hand sides, face values and runs (tests/golden/*_exp.npz)
and against the oracle (which reproduces those fixtures bit for bit on the CPU).  The
device evaluates WENO-EXP's pow(l2h2, 3) as a once-rounded cube and THINC's exp, log and
tanh with the device math library, where glibc may differ in the last bit, so the
comparisons carry a tolerance: 1e-12 relative to the largest magnitude of each field for a
right-hand side or a face value, 1e-10 for a state after a run, 1e-12 relative for each dt.
The batched solver must equal the single-domain solver bit for bit, and the mixed policy
must stay within 1e-5 relative L1 of the FP64 one on a smooth case."""
import numpy as np
import pytest

import oracle as orc
import wenofv
from conftest import golden

pytestmark = pytest.mark.gpu

RHS = golden("solver_rhs_exp.npz")
RUNS = golden("solver_runs_exp.npz")
CASE = {
    0: ((0.0, 1.0, 0.0, 1.0), 1.2, 1.0, 1.0, 1.0, 0.0),
    1: ((-1.0, 1.0, -1.0, 1.0), 1000.0, 1.0, 30.0, 30.0, 0.0),
    2: ((0.0, 1.0, 0.0, 1.0), 1000.0, 1.0, 70.0, 70.0, 0.0),
}


def close(got, ref, rel, axis_fields=0):
    """|got - ref| <= rel * max|ref| per field (fields along `axis_fields`)."""
    g = np.moveaxis(got, axis_fields, 0).reshape(got.shape[axis_fields], -1)
    r = np.moveaxis(ref, axis_fields, 0).reshape(ref.shape[axis_fields], -1)
    for k in range(r.shape[0]):
        scale = max(np.abs(r[k]).max(), 1e-300)
        err = np.abs(g[k] - r[k]).max()
        assert err <= rel * scale, (k, err, scale)


def rhs_tags():
    return sorted(k[:-4] for k in RHS.files if k.endswith("_rhs"))


@pytest.mark.parametrize("tag", rhs_tags())
def test_rhs_and_faces_match_reference(tag):
    kind, s, r, f, b = tag.split("_")
    p = RHS[kind + "_params"]
    U = RHS[kind + "_U"]
    ny, nx = U.shape[1:]
    print(f"exp case {tag} {nx}x{ny}", flush=True)
    sol = wenofv.Solver(nx, ny, rho10=p[0], rho20=p[1], sound1=p[2], sound2=p[3], preb=p[4], scheme=int(s[1]),
                        recon=int(r[1]), flux=int(f[1]), bp=int(b[1]))
    sol.set_state(U)
    rhs, faces, theta = sol.assemble_rhs(want_faces=True)
    close(rhs, RHS[tag + "_rhs"], 1e-12)
    if tag + "_faces" in RHS.files:
        # faces [ny][nx][side][var][gp]: compare per variable
        close(faces.reshape(ny, nx, 4, 4, 4), RHS[tag + "_faces"].reshape(ny, nx, 4, 4, 4), 1e-12, axis_fields=3)
    if int(b[1]):
        np.testing.assert_array_equal((theta < 1 - 1e-8).astype(float), RHS[tag + "_thflag"])


def run_names():
    return sorted(k[:-5] for k in RUNS.files if k.endswith("_opts"))


@pytest.mark.parametrize("name", run_names())
def test_time_loop_matches_reference(name):
    tc, nx, ny, rk, cfl, scheme, bp, flux, recon, fixed, tend = RUNS[name + "_opts"]
    dom, rho10, rho20, c1, c2, preb = CASE[int(tc)]
    print(f"exp run {name}", flush=True)
    sol = wenofv.Solver(int(nx), int(ny), domain=dom, rho10=rho10, rho20=rho20, sound1=c1, sound2=c2, preb=preb,
                        scheme=int(scheme), recon=int(recon), flux=int(flux), bp=int(bp), rk=int(rk), cfl=cfl)
    ref_dts = RUNS[name + "_dts"]
    sol.set_state(RUNS[name + "_init"])
    if fixed:
        n, t, dts = sol.advance(fixed_steps=int(tend), max_dts=len(ref_dts) + 4)
    else:
        n, t, dts = sol.advance(t_end=tend, max_dts=len(ref_dts) + 4)
    assert n == len(ref_dts)
    np.testing.assert_allclose(dts, ref_dts, rtol=1e-12)
    close(sol.get_state(), RUNS[name + "_final"], 1e-10)
    assert sol.flags() == 0


def state(rng, nx, ny, params, sharp):
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


COMBOS = [(2, r, f, b) for r in (0, 1, 2, 3) for f in (0, 1) for b in (0, 1)] + \
         [(1, 4, f, b) for f in (0, 1) for b in (0, 1)]


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("combo", COMBOS)
def test_rhs_matches_oracle_on_ragged_grids(combo, bc):
    s, r, f, b = combo
    rng = np.random.default_rng(1000 + 97 * r + 13 * f + 5 * b + bc)
    nx, ny = (70, 37) if bc else (64, 45)   # strips of 56 columns: one full and one partial
    params = (1.0, 0.1, 10.0, 10.0, 0.0)
    U = state(rng, nx, ny, params, bool(b) and r < 2)
    print(f"exp oracle rhs s{s} r{r} f{f} b{b} bc{bc} {nx}x{ny}", flush=True)
    kw = dict(scheme=s, recon=r, flux=f, bp=b, rho10=params[0], rho20=params[1], sound1=params[2],
              sound2=params[3], preb=params[4], domain=(0.0, 2.0, 0.0, 1.0))
    o = orc.Oracle(nx, ny, bc=bc, nranks=1, **kw)
    sol = wenofv.Solver(nx, ny, bc=bc, geometry_ranks=1, rows_per_strip=8, **kw)
    sol.set_state(U)
    rhs, faces, theta = sol.assemble_rhs(want_faces=True)
    r_o, f_o, t_o, err = o.assemble_rhs(U, want_faces=True)
    close(rhs, r_o, 1e-12)
    close(faces.reshape(ny, nx, 4, 4, 4), f_o.reshape(ny, nx, 4, 4, 4), 1e-12, axis_fields=3)
    np.testing.assert_allclose(theta, t_o, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme,recon", [(2, 0), (2, 1), (1, 4)])
def test_batch_equals_single_solver(precision, scheme, recon):
    rng = np.random.default_rng(5 + recon)
    params = (1.0, 0.1, 10.0, 10.0, 0.0)
    shapes = [(64, 40), (90, 33), (57, 70)]
    opts = dict(scheme=scheme, recon=recon, flux=1, bp=1, rk=2, cfl=0.4, precision=precision)
    insts, states = [], []
    for k, (nx, ny) in enumerate(shapes):
        insts.append(dict(nx=nx, ny=ny, bc=k % 2, domain=(0.0, 1.0, 0.0, 1.0), rho10=params[0], rho20=params[1],
                          sound1=params[2], sound2=params[3], preb=params[4]))
        states.append(state(rng, nx, ny, params, False))
    print(f"exp batch p{precision} s{scheme} r{recon}", flush=True)
    b = wenofv.Batch(insts, geometry_ranks=1, **opts)
    for k, U in enumerate(states):
        b.set_state(k, U)
    n, t, f = b.advance(fixed_steps=5)
    assert (n == 5).all() and (f == 0).all()
    for k, (nx, ny) in enumerate(shapes):
        d = dict(insts[k])
        sol = wenofv.Solver(nx, ny, bc=d.pop("bc"), geometry_ranks=1, **{x: d[x] for x in d if x not in ("nx", "ny")},
                            **opts)
        sol.set_state(states[k])
        sol.advance(fixed_steps=5)
        np.testing.assert_array_equal(b.get_state(k), sol.get_state())


@pytest.mark.parametrize("scheme,recon", [(2, 0), (2, 2), (1, 4)])
def test_mixed_matches_fp64(scheme, recon):
    rng = np.random.default_rng(11)
    nx, ny = 96, 64
    params = (1.0, 0.1, 10.0, 10.0, 0.0)
    U = state(rng, nx, ny, params, False)
    kw = dict(scheme=scheme, recon=recon, flux=1, bp=1, rk=2, cfl=0.4, rho10=params[0], rho20=params[1],
              sound1=params[2], sound2=params[3], preb=params[4], geometry_ranks=1)
    out = []
    for prec in (0, 1):
        sol = wenofv.Solver(nx, ny, precision=prec, **kw)
        sol.set_state(U)
        sol.advance(fixed_steps=20)
        out.append(sol.get_state())
    for v in range(4):
        rel = np.abs(out[1][v] - out[0][v]).sum() / np.abs(out[0][v]).sum()
        assert rel <= 1e-5, (v, rel)
