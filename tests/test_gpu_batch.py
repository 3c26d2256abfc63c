"""Batched solver: every instance of a batch reproduces a single-domain solver with the
same options bit for bit (tolerance 0), in every precision, with uneven grids,
periodic and wall boundaries and per-instance time steps; in all-FP64 every instance also
equals the oracle run on its own (state and final time, bit for bit)."""
import numpy as np
import pytest

import cases
import oracle as orc
import wenofv

pytestmark = pytest.mark.gpu


def make_instances(seed, n):
    specs = cases.batch_specs(n=n, seed=seed)
    out = []
    for k, sp in enumerate(specs):
        sp = dict(sp, nx=[64, 70, 96, 128][k % 4], ny=[64, 37, 80, 50][(k + 1) % 4])
        wl = cases.batch_instance(sp)
        out.append(wl)
    return out


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
@pytest.mark.parametrize("rk,recon", [(2, 0), (1, 1), (0, 3)])
def test_batch_instances_equal_single_solvers(precision, rk, recon):
    wls = make_instances(11 + rk, 5)
    opts = dict(scheme=1, recon=recon, flux=1, bp=1, rk=rk, cfl=0.4, precision=precision)
    insts = []
    for k, wl in enumerate(wls):
        d = dict(nx=wl.nx, ny=wl.ny, bc=1 if k % 2 == 0 else 0, domain=wl.domain, **wl.params)
        insts.append(d)
    print(f"batch precision{precision} rk{rk} recon{recon}", flush=True)
    b = wenofv.Batch(insts, geometry_ranks=1, **opts)
    for k, wl in enumerate(wls):
        b.set_state(k, wl.U0)
    n, t, f = b.advance(fixed_steps=7)
    assert (n == 7).all() and (f == 0).all()
    for k, wl in enumerate(wls):
        sol = wenofv.Solver(wl.nx, wl.ny, domain=wl.domain, bc=insts[k]["bc"], geometry_ranks=1, **wl.params,
                            **opts)
        sol.set_state(wl.U0)
        ns, ts, _ = sol.advance(fixed_steps=7)
        assert ts == t[k]
        np.testing.assert_array_equal(b.get_state(k), sol.get_state())
        if precision == 0:
            o = orc.Oracle(wl.nx, wl.ny, domain=wl.domain, bc=insts[k]["bc"], nranks=1, **wl.params,
                           **{q: opts[q] for q in ("scheme", "recon", "flux", "bp", "rk", "cfl")})
            U_cpu, _, t_cpu, err = o.run(wl.U0, fixed_steps=7)
            assert err == 0 and t[k] == t_cpu
            np.testing.assert_array_equal(b.get_state(k), U_cpu)


def test_batch_time_driven_per_instance_dt():
    wls = make_instances(3, 4)
    insts = [dict(nx=wl.nx, ny=wl.ny, bc=1, domain=wl.domain, **wl.params) for wl in wls]
    opts = dict(scheme=0, recon=0, flux=0, bp=1, rk=2, cfl=0.4, precision=0)
    b = wenofv.Batch(insts, **opts)
    for k, wl in enumerate(wls):
        b.set_state(k, wl.U0)
    n, t, f = b.advance(t_end=2e-3)
    assert np.allclose(t, 2e-3, atol=1e-10) and (f == 0).all()
    for k, wl in enumerate(wls):
        sol = wenofv.Solver(wl.nx, wl.ny, domain=wl.domain, bc=1, geometry_ranks=1, **wl.params, **opts)
        sol.set_state(wl.U0)
        ns, ts, _ = sol.advance(t_end=2e-3)
        assert ns == n[k] and ts == t[k]
        np.testing.assert_array_equal(b.get_state(k), sol.get_state())
        o = orc.Oracle(wl.nx, wl.ny, domain=wl.domain, bc=1, nranks=1, **wl.params,
                       **{q: opts[q] for q in ("scheme", "recon", "flux", "bp", "rk", "cfl")})
        U_cpu, dts_cpu, t_cpu, err = o.run(wl.U0, t_end=2e-3)
        assert err == 0 and len(dts_cpu) == n[k] and t_cpu == t[k]
        np.testing.assert_array_equal(b.get_state(k), U_cpu)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme,recon", [(2, 0), (1, 4)])
def test_batch_seven_point_forms_equal_single_solvers(precision, scheme, recon):
    """WENO-EXP and primitive3/THINC (halo 4, 56-column strips) in a batch of uneven
    grids: every instance equals its single-domain solver bit for bit."""
    wls = make_instances(21 + recon, 4)
    opts = dict(scheme=scheme, recon=recon, flux=1, bp=1, rk=2, cfl=0.4, precision=precision)
    insts = [dict(nx=wl.nx, ny=wl.ny, bc=1 if k % 2 == 0 else 0, domain=wl.domain, **wl.params)
             for k, wl in enumerate(wls)]
    print(f"batch 7-point precision{precision} scheme{scheme} recon{recon}", flush=True)
    b = wenofv.Batch(insts, geometry_ranks=1, **opts)
    for k, wl in enumerate(wls):
        b.set_state(k, wl.U0)
    n, t, f = b.advance(fixed_steps=5)
    assert (n == 5).all() and (f == 0).all()
    for k, wl in enumerate(wls):
        sol = wenofv.Solver(wl.nx, wl.ny, domain=wl.domain, bc=insts[k]["bc"], geometry_ranks=1, **wl.params,
                            **opts)
        sol.set_state(wl.U0)
        sol.advance(fixed_steps=5)
        np.testing.assert_array_equal(b.get_state(k), sol.get_state())
