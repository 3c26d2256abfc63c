"""The benchmark configurations C2-C5 on the GPU (C1 runs in full against the reference's own
solver in test_gpu_full_runs.py).

- C2, the 1024 x 512 walled liquid column hit by a ratio-2 pressure step (WENO-Z, HLLC,
  limiter, RK3, CFL 0.4): the all-FP64 solver equals the oracle bit for bit (every dt and the
  state) on the same case at 128 x 64 over 100 steps; at full size over its 2000 steps, in
  both precisions, the state stays finite with no error flag, z1 in [0, 1] and per-phase mass
  conserved.
- C3, the 4096 x 2048 column with a 16-mode radius perturbation, 200 steps: mixed against
  all-FP64 within 1e-5 relative L1 on z1, rho, u, v, with the same properties.
- C4, the 256 seeded instances with grids {64..512}^2, 500 steps in one batch, mixed and
  all-FP64: the batch within 1e-5 relative L1 on z1, rho, u, v (precision 2, whose
  reconstruction is FP64 too, also on the conserved fields), every instance with the
  properties; and one all-FP64 batch of an instance of every grid size equal to the oracle
  bit for bit over 5 steps at full size.
- C5, 8192 x 8192 with 64 columns and a ratio-3 step, 20 steps: mixed against all-FP64 within
  1e-5, with the properties.
Relative L1 of a velocity component is taken against the velocity magnitude's L1 norm."""
import numpy as np
import pytest

import cases
import oracle as orc
import wenofv

pytestmark = pytest.mark.gpu


def derived(U, rho10, rho20, sound1, sound2, **_):
    """z1 (include/Equations.hpp:84-99), rho, u, v of a state."""
    c1, c2 = sound1 ** 2, sound2 ** 2
    q = rho20 * c2 - rho10 * c1
    qs = U[1] * c2 - U[0] * c1
    r = ((q - qs) + np.sqrt((q - qs) ** 2 + 4.0 * U[0] * c1 * U[1] * c2)) / (2.0 * U[1] * c2)
    rho = U[0] + U[1]
    return r / (1 + r), rho, U[2] / rho, U[3] / rho


def gate_errors(Um, Ud, params):
    """Relative L1 of z1, rho, u, v of Um against Ud: (numerators, denominators)."""
    z1m, rm, um, vm = derived(Um, **params)
    z1d, rd, ud, vd = derived(Ud, **params)
    vel = np.hypot(ud, vd).sum()
    num = {"z1": np.abs(z1m - z1d).sum(), "rho": np.abs(rm - rd).sum(), "u": np.abs(um - ud).sum(),
           "v": np.abs(vm - vd).sum()}
    den = {"z1": np.abs(z1d).sum(), "rho": np.abs(rd).sum(), "u": vel, "v": vel}
    return num, den


def check_properties(U, U0, params, mass_tol):
    """Finite state, z1 in [0, 1], per-phase mass (walls: nothing enters or leaves)."""
    assert np.isfinite(U).all()
    z1 = derived(U, **params)[0]
    assert z1.min() >= 0.0 and z1.max() <= 1.0, (z1.min(), z1.max())
    m0, m = U0[:2].sum(axis=(1, 2)), U[:2].sum(axis=(1, 2))
    drift = np.abs(m / m0 - 1).max()
    assert drift <= mass_tol, drift
    return drift


def run_solver(wl, steps, precision):
    sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs(precision=precision))
    sol.set_state(wl.U0)
    n, t, _ = sol.advance(fixed_steps=steps)
    assert n == steps and sol.flags() == 0
    U = sol.get_state()
    sol.close()
    return U


MASS_TOL = {0: 1e-12, 1: 1e-8}


def test_c2_liquid_column_matches_oracle_bitwise():
    wl = cases.liquid_column(nx=128, ny=64, precision=0)
    print("C2 128x64 FP64 vs oracle, 100 steps", flush=True)
    sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    sol.set_state(wl.U0)
    n, t, dts = sol.advance(fixed_steps=100, max_dts=100)
    o = orc.Oracle(wl.nx, wl.ny, domain=wl.domain, bc=wl.bc, nranks=1, **wl.params,
                   **{k: wl.opts[k] for k in ("scheme", "recon", "flux", "bp", "rk", "cfl")})
    U_cpu, dts_cpu, t_cpu, err = o.run(wl.U0, fixed_steps=100)
    assert n == 100 and err == 0
    np.testing.assert_array_equal(dts, dts_cpu)
    assert t == t_cpu
    np.testing.assert_array_equal(sol.get_state(), U_cpu)


@pytest.mark.parametrize("precision", [0, 1])
def test_c2_full_size_properties(precision):
    wl = cases.liquid_column(precision=precision)
    print(f"C2 1024x512 precision{precision}, {wl.steps} steps", flush=True)
    U = run_solver(wl, wl.steps, precision)
    drift = check_properties(U, wl.U0, wl.params, MASS_TOL[precision])
    print(f"C2 precision{precision}: mass drift {drift:.2e}", flush=True)
    assert not np.array_equal(U, wl.U0)


@pytest.fixture(scope="module")
def c3():
    return cases.perturbed_column(precision=1)


def test_c3_perturbed_column_mixed_against_fp64(c3):
    print(f"C3 4096x2048 mixed and FP64, {c3.steps} steps", flush=True)
    Um = run_solver(c3, c3.steps, 1)
    Ud = run_solver(c3, c3.steps, 0)
    num, den = gate_errors(Um, Ud, c3.params)
    errs = {k: num[k] / den[k] for k in num}
    print(f"C3 mixed vs FP64 rel L1: {errs}", flush=True)
    assert max(errs.values()) <= 1e-5, errs
    check_properties(Um, c3.U0, c3.params, MASS_TOL[1])
    check_properties(Ud, c3.U0, c3.params, MASS_TOL[0])


@pytest.fixture(scope="module")
def c4():
    return [cases.batch_instance(sp) for sp in cases.batch_specs(256, seed=4)]


def run_batch(wls, steps, precision, idx=None):
    idx = range(len(wls)) if idx is None else idx
    sel = [wls[k] for k in idx]
    insts = [dict(nx=w.nx, ny=w.ny, bc=w.bc, domain=w.domain, **w.params) for w in sel]
    b = wenofv.Batch(insts, **dict(sel[0].opts, precision=precision))
    for k, w in enumerate(sel):
        b.set_state(k, w.U0)
    n, t, f = b.advance(fixed_steps=steps)
    assert (n == steps).all() and (f == 0).all()
    out = [b.get_state(k) for k in range(len(sel))], t
    b.close()
    return out


@pytest.fixture(scope="module")
def c4_fp64(c4):
    S, _ = run_batch(c4, c4[0].steps, 0)
    for w, Ud in zip(c4, S):
        check_properties(Ud, w.U0, w.params, MASS_TOL[0])
    return S


@pytest.mark.parametrize("precision", [1, 2])
def test_c4_batch256_mixed_against_fp64(c4, c4_fp64, precision):
    """The batch within 1e-5 relative L1 of all-FP64 on z1, rho, u, v; precision 2 (FP64
    reconstruction) also on the conserved fields (z1 rho1, z2 rho2, rho u, rho v)."""
    steps = c4[0].steps
    print(f"C4 batch256 precision {precision} and FP64, {steps} steps", flush=True)
    Sm, _ = run_batch(c4, steps, precision)
    keys = ("z1", "rho", "u", "v", "z1rho1", "z2rho2", "rhou", "rhov")
    num = {k: 0.0 for k in keys}
    den = dict(num)
    for w, Um, Ud in zip(c4, Sm, c4_fp64):
        check_properties(Um, w.U0, w.params, MASS_TOL[1])
        n_, d_ = gate_errors(Um, Ud, w.params)
        mom = np.hypot(Ud[2], Ud[3]).sum()
        for r, k in enumerate(("z1rho1", "z2rho2", "rhou", "rhov")):
            n_[k] = np.abs(Um[r] - Ud[r]).sum()
            d_[k] = mom if r >= 2 else np.abs(Ud[r]).sum()
        for k in keys:
            num[k] += n_[k]
            den[k] += d_[k]
    errs = {k: num[k] / den[k] for k in keys}
    print(f"C4 precision {precision} vs FP64 rel L1 over the batch: {errs}", flush=True)
    gate = keys if precision == 2 else keys[:4]
    assert max(errs[k] for k in gate) <= 1e-5, errs


def test_c4_every_grid_size_matches_oracle_bitwise(c4):
    """One instance of each of the 16 grid sizes, batched in all-FP64, against the oracle."""
    first = {}
    for k, w in enumerate(c4):
        first.setdefault((w.nx, w.ny), k)
    idx = sorted(first.values())
    assert len(idx) == 16
    print("C4 16 grid sizes FP64 batch vs oracle, 5 steps", flush=True)
    S, t = run_batch(c4, 5, 0, idx)
    for j, k in enumerate(idx):
        w = c4[k]
        o = orc.Oracle(w.nx, w.ny, domain=w.domain, bc=w.bc, nranks=1, **w.params,
                       **{q: w.opts[q] for q in ("scheme", "recon", "flux", "bp", "rk", "cfl")})
        U_cpu, dts, t_cpu, err = o.run(w.U0, fixed_steps=5)
        assert err == 0 and t[j] == t_cpu, (w.nx, w.ny)
        np.testing.assert_array_equal(S[j], U_cpu, err_msg=f"instance {k} ({w.nx}x{w.ny})")


def test_c5_multi_column_mixed_against_fp64():
    wl = cases.multi_column(precision=1)
    steps = 20
    print(f"C5 8192x8192 mixed and FP64, {steps} steps", flush=True)
    Um = run_solver(wl, steps, 1)
    check_properties(Um, wl.U0, wl.params, MASS_TOL[1])
    Ud = run_solver(wl, steps, 0)
    num, den = gate_errors(Um, Ud, wl.params)
    errs = {k: num[k] / den[k] for k in num}
    print(f"C5 mixed vs FP64 rel L1: {errs}", flush=True)
    assert max(errs.values()) <= 1e-5, errs
    check_properties(Ud, wl.U0, wl.params, MASS_TOL[0])
