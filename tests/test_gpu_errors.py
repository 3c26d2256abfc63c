"""Error behaviour of the C ABI (include/wenofv.h): invalid options are refused at creation
with WFV_ERR_CONFIG and a message naming the option (the reference's std::stoi throw /
MPI_Abort, src/WENOFV.cpp:43-63); null pointers and bad indices return WFV_ERR_ARG; a state
that yields a NaN time step stops the run with WFV_ERR_NONFINITE_DT and the reference's
message (src/WENOFV.cpp:502-531), after which the context (or batch) still works; two
contexts on one GPU do not interact.  Invalid inputs are all caught on the host; the NaN
state is ordinary data to the kernels."""
import ctypes

import numpy as np
import pytest

import wenofv

pytestmark = pytest.mark.gpu

PRM = dict(rho10=1.0, rho20=0.1, sound1=10.0, sound2=10.0, preb=0.0)


def state(nx, ny):
    X, Y = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny)
    z1 = 0.5 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    p = np.ones_like(X)
    a, b = z1 * (1.0 + p / 100), (1 - z1) * (0.1 + p / 100)
    return np.stack([a, b, (a + b) * 0.3, (a + b) * -0.2])


@pytest.mark.parametrize("bad,msg", [
    (dict(nx=2), "at least 3 cells"),
    (dict(nx=0), "must be positive"),
    (dict(scheme=3), "SCHEME"),
    (dict(recon=5), "RECONSTRUCTIONTYPE"),
    (dict(flux=2), "RIEMANNFLUX"),
    (dict(rk=3), "RKMETH"),
    (dict(precision=4), "precision"),
    (dict(precision=3, scheme=2), "5-point"),
    (dict(bc=4), "boundary"),
    (dict(cfl=0.0), "CFL"),
    (dict(nx=70, geometry_ranks=4), "divisible"),
    (dict(domain=(1.0, 1.0, 0.0, 1.0)), "empty domain"),
])
def test_invalid_options_are_refused_at_creation(bad, msg):
    kw = dict(nx=16, ny=16, domain=(0.0, 1.0, 0.0, 1.0), geometry_ranks=1, **PRM)
    kw.update(bad)
    nx, ny = kw.pop("nx"), kw.pop("ny")
    with pytest.raises(wenofv.WenofvError) as e:
        wenofv.Solver(nx, ny, **kw)
    assert f"({wenofv.ERR_CONFIG})" in str(e.value) and msg in str(e.value), str(e.value)


def test_null_pointers_and_bad_arguments():
    L = wenofv.lib()
    sol = wenofv.Solver(16, 16, geometry_ranks=1, **PRM)
    h = sol._h
    assert L.wfv_set_state(h, None) == wenofv.ERR_ARG
    assert L.wfv_get_state(h, None) == wenofv.ERR_ARG
    assert L.wfv_compute_dt(h, None) == wenofv.ERR_ARG
    assert L.wfv_autotune(h, 0, None) == wenofv.ERR_ARG
    assert L.wfv_set_rows_per_strip(h, -1) == wenofv.ERR_ARG
    out = ctypes.c_void_p()
    assert L.wfv_create(None, ctypes.byref(sol.case), ctypes.byref(out)) == wenofv.ERR_ARG
    assert L.wfv_create(ctypes.byref(sol.opt), None, ctypes.byref(out)) == wenofv.ERR_ARG
    # the context is untouched by the refused calls
    sol.set_state(state(16, 16))
    n, t, _ = sol.advance(fixed_steps=2)
    assert n == 2 and t > 0 and sol.flags() == 0


@pytest.mark.parametrize("precision", [0, 1])
def test_nan_state_stops_the_run_and_the_context_recovers(precision):
    sol = wenofv.Solver(32, 24, bc=1, geometry_ranks=1, scheme=1, flux=1, bp=1, precision=precision, **PRM)
    U = state(32, 24)
    bad = U.copy()
    bad[0, 5, 7] = np.nan
    print(f"nan state precision{precision}", flush=True)
    sol.set_state(bad)
    with pytest.raises(wenofv.WenofvError) as e:
        sol.advance(fixed_steps=3)
    assert f"error {wenofv.ERR_NONFINITE_DT}" in str(e.value) and "nan DeltaT" in str(e.value), str(e.value)
    # a valid state runs again on the same context
    sol.set_state(U)
    n, t, _ = sol.advance(fixed_steps=3)
    assert n == 3 and t > 0 and np.isfinite(sol.get_state()).all()


def test_batch_arguments():
    inst = [dict(nx=16, ny=16, bc=0, domain=(0.0, 1.0, 0.0, 1.0), **PRM)]
    with pytest.raises(wenofv.WenofvError):
        wenofv.Batch([])
    with pytest.raises(wenofv.WenofvError):
        wenofv.Batch([dict(inst[0], nx=2)])
    b = wenofv.Batch(inst)
    BL = wenofv._batch_lib()
    U = np.ascontiguousarray(state(16, 16))
    assert BL.wfv_batch_set_state(b._h, 1, U.ctypes.data) == wenofv.ERR_ARG
    assert BL.wfv_batch_set_state(b._h, -1, U.ctypes.data) == wenofv.ERR_ARG
    assert BL.wfv_batch_get_state(b._h, 0, None) == wenofv.ERR_ARG
    with pytest.raises(wenofv.WenofvError):
        b.set_ghost_cells(0, np.zeros((4, 24, 24)))
    b.set_state(0, state(16, 16))
    n, t, f = b.advance(fixed_steps=2)
    assert (n == 2).all() and (f == 0).all()


def test_batch_recovers_after_a_nan_instance():
    inst = [dict(nx=16, ny=16, bc=0, domain=(0.0, 1.0, 0.0, 1.0), **PRM),
            dict(nx=24, ny=20, bc=1, domain=(0.0, 1.0, 0.0, 1.0), **PRM)]
    b = wenofv.Batch(inst, scheme=1, flux=1, bp=1)
    good = [state(16, 16), state(24, 20)]
    bad = good[1].copy()
    bad[2, 3, 4] = np.nan
    b.set_state(0, good[0])
    b.set_state(1, bad)
    print("batch nan instance", flush=True)
    with pytest.raises(wenofv.WenofvError):
        b.advance(fixed_steps=2)
    b.set_state(0, good[0])
    b.set_state(1, good[1])
    n, t, f = b.advance(fixed_steps=2)
    assert (n == 2).all() and np.isfinite(b.get_state(1)).all()


def test_contexts_are_independent():
    """Two solvers on one GPU, advanced in alternation, each equal to a solver run alone
    (one stream per context, nothing shared)."""
    opts = dict(geometry_ranks=1, scheme=1, flux=1, bp=1, **PRM)
    a = wenofv.Solver(32, 24, bc=1, **opts)
    b = wenofv.Solver(20, 28, bc=0, precision=1, **opts)
    Ua, Ub = state(32, 24), state(20, 28)[:, ::-1].copy()
    a.set_state(Ua)
    b.set_state(Ub)
    for _ in range(3):
        a.advance(fixed_steps=2)
        b.advance(fixed_steps=3)
    solo_a = wenofv.Solver(32, 24, bc=1, **opts)
    solo_a.set_state(Ua)
    for _ in range(3):
        solo_a.advance(fixed_steps=2)
    solo_b = wenofv.Solver(20, 28, bc=0, precision=1, **opts)
    solo_b.set_state(Ub)
    for _ in range(3):
        solo_b.advance(fixed_steps=3)
    np.testing.assert_array_equal(a.get_state(), solo_a.get_state())
    np.testing.assert_array_equal(b.get_state(), solo_b.get_state())


@pytest.mark.parametrize("precision", [0, 1])
def test_rk1_run_of_no_steps_leaves_the_state_unchanged(precision):
    """SSP-RK1 writes its stage into a second register and copies it back; a run that takes
    no step (t_end = t_start, or a state whose first time step is NaN) must leave the state
    as it was, as the reference's loop does when it takes no step (src/WENOFV.cpp:318-332)."""
    nx, ny = 40, 24
    U0 = state(nx, ny)
    sol = wenofv.Solver(nx, ny, rk=0, precision=precision, bc=1, geometry_ranks=1, scheme=1, flux=1, bp=1, **PRM)
    print(f"rk1 zero steps precision{precision}", flush=True)
    sol.set_state(U0)
    n, t, _ = sol.advance(t_end=0.0, t_start=0.0)
    assert n == 0 and t == 0.0
    np.testing.assert_array_equal(sol.get_state(), U0)
    # a NaN first time step: the run stops before its first step, the state stays
    bad = U0.copy()
    bad[1, 3, 5] = np.nan
    sol.set_state(bad)
    with pytest.raises(wenofv.WenofvError):
        sol.advance(fixed_steps=2)
    assert np.array_equal(sol.get_state(), bad, equal_nan=True)
    # and a real step still changes it
    sol.set_state(U0)
    n, t, _ = sol.advance(fixed_steps=1)
    assert n == 1 and not np.array_equal(sol.get_state(), U0)


def test_rk1_batch_run_of_no_steps_leaves_the_states_unchanged():
    inst = [dict(nx=16, ny=16, bc=0, domain=(0.0, 1.0, 0.0, 1.0), **PRM),
            dict(nx=24, ny=20, bc=1, domain=(0.0, 1.0, 0.0, 1.0), **PRM)]
    b = wenofv.Batch(inst, rk=0, scheme=1, flux=1, bp=1)
    good = [state(16, 16), state(24, 20)]
    for k in range(2):
        b.set_state(k, good[k])
    print("rk1 batch zero steps", flush=True)
    n, t, f = b.advance(t_end=0.0)
    assert (n == 0).all()
    for k in range(2):
        np.testing.assert_array_equal(b.get_state(k), good[k])
    # one instance stops at a NaN time step before its first step; the other runs
    bad = good[1].copy()
    bad[0, 2, 2] = np.nan
    b.set_state(1, bad)
    with pytest.raises(wenofv.WenofvError):
        b.advance(fixed_steps=2)
    assert np.array_equal(b.get_state(1), bad, equal_nan=True)
    assert not np.array_equal(b.get_state(0), good[0])


def test_rows_per_strip_beyond_the_grid_is_the_whole_grid():
    L = wenofv.lib()
    sol = wenofv.Solver(40, 24, geometry_ranks=1, **PRM)
    assert L.wfv_set_rows_per_strip(sol._h, 2 ** 31 - 1) == wenofv.OK
    assert sol.rows_per_strip() == 24
    sol.set_state(state(40, 24))
    n, t, _ = sol.advance(fixed_steps=2)
    assert n == 2 and sol.flags() == 0
    assert L.wfv_set_rows_per_strip(sol._h, 0) == wenofv.OK  # 0: the default, as for a batch
    assert 1 <= sol.rows_per_strip() <= 24
    inst = [dict(nx=16, ny=16, bc=0, domain=(0.0, 1.0, 0.0, 1.0), **PRM)]
    b = wenofv.Batch(inst, rows_per_strip=1000)
    assert b.rows_per_strip(0) == 16
    b.set_rows_per_strip(0, 2 ** 31 - 1)
    assert b.rows_per_strip(0) == 16
