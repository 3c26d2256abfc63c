"""Launch-shape autotuner: tuning leaves the state untouched, changes only the launch shape
(results stay bit-identical to the untuned solver, tolerance 0), and the cache of earlier
tunings restores the same choice."""
import numpy as np
import pytest

import cases
import wenofv

pytestmark = pytest.mark.gpu


def small_batch(n=6):
    specs = cases.batch_specs(n=n, seed=21)
    wls = []
    for k, sp in enumerate(specs):
        sp = dict(sp, nx=[64, 128, 96][k % 3], ny=[64, 48, 128][(k + 1) % 3])
        wls.append(cases.batch_instance(sp))
    return wls


@pytest.mark.parametrize("precision", [0, 1])
def test_solver_autotune_keeps_results(precision, tmp_path, monkeypatch):
    monkeypatch.setenv("WENOFV_TUNING_CACHE", str(tmp_path / "tune.json"))
    wl = cases.batch_instance(dict(cases.batch_specs(1, seed=5)[0], nx=200, ny=150), precision=precision)
    print(f"solver autotune precision{precision}", flush=True)
    ref = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    ref.set_state(wl.U0)
    ref.advance(fixed_steps=5)
    sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    sol.set_state(wl.U0)
    res = sol.autotune(trials=1)
    assert res["rows_per_strip"] == sol.rows_per_strip() and res["stage_ms"] > 0
    np.testing.assert_array_equal(sol.get_state(), wl.U0)  # tuning leaves the state alone
    sol.advance(fixed_steps=5)
    np.testing.assert_array_equal(sol.get_state(), ref.get_state())
    # the cache gives the same choice without timing
    again = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    assert again.autotune()["rows_per_strip"] == res["rows_per_strip"]
    assert again.rows_per_strip() == res["rows_per_strip"]


def test_solver_autotune_on_a_medium_grid_keeps_results(tmp_path, monkeypatch):
    """1024 x 512: the tuner adds its occupancy candidates (strip heights whose wavefront
    count fills whole rounds of the device; on 256 CUs 2, 3, 4, 5 and 10 rows, which
    tests/strip_plan_check.cpp checks on the host) to the fixed list."""
    monkeypatch.setenv("WENOFV_TUNING_CACHE", str(tmp_path / "tune.json"))
    wl = cases.liquid_column(nx=1024, ny=512, precision=0)
    print("solver autotune 1024x512", flush=True)
    ref = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    ref.set_state(wl.U0)
    ref.advance(fixed_steps=3)
    sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    sol.set_state(wl.U0)
    res = sol.autotune(trials=1, cache=False)
    print(res, flush=True)
    np.testing.assert_array_equal(sol.get_state(), wl.U0)
    sol.advance(fixed_steps=3)
    np.testing.assert_array_equal(sol.get_state(), ref.get_state())


@pytest.mark.parametrize("rows", [1, 2, 5, 8, 33, 200])
def test_solver_rows_per_strip_is_invisible(rows):
    wl = cases.batch_instance(dict(cases.batch_specs(1, seed=6)[0], nx=90, ny=77), precision=0)
    a = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    b = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs(rows_per_strip=rows))
    for s in (a, b):
        s.set_state(wl.U0)
        s.advance(fixed_steps=4)
    np.testing.assert_array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_autotune_keeps_results(precision, tmp_path, monkeypatch):
    monkeypatch.setenv("WENOFV_TUNING_CACHE", str(tmp_path / "tune.json"))
    wls = small_batch()
    insts = [dict(nx=w.nx, ny=w.ny, bc=1, domain=w.domain, **w.params) for w in wls]
    opts = dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=precision)
    print(f"batch autotune precision{precision}", flush=True)
    ref = wenofv.Batch(insts, **opts)
    tun = wenofv.Batch(insts, **opts)
    for k, w in enumerate(wls):
        ref.set_state(k, w.U0)
        tun.set_state(k, w.U0)
    tun.advance(fixed_steps=2)  # the tuner must also work on a batch whose loop has finished
    for k, w in enumerate(wls):
        tun.set_state(k, w.U0)
    res = tun.autotune(trials=1)
    assert set(res["rows_per_strip"]) == {f"{w.nx}x{w.ny}" for w in wls}
    assert res["persistent"] == tun.schedule()
    for k, w in enumerate(wls):
        assert tun.rows_per_strip(k) == res["rows_per_strip"][f"{w.nx}x{w.ny}"]
        np.testing.assert_array_equal(tun.get_state(k), w.U0)
    n1, t1, f1 = ref.advance(fixed_steps=6)
    n2, t2, f2 = tun.advance(fixed_steps=6)
    assert (n1 == n2).all() and (t1 == t2).all() and (f2 == 0).all()
    for k in range(len(wls)):
        np.testing.assert_array_equal(tun.get_state(k), ref.get_state(k))
    cached = wenofv.Batch(insts, **opts).autotune()
    assert cached["rows_per_strip"] == res["rows_per_strip"] and cached["persistent"] == res["persistent"]


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_schedule_is_invisible(persistent):
    wls = small_batch(4)
    insts = [dict(nx=w.nx, ny=w.ny, bc=0, domain=w.domain, **w.params) for w in wls]
    opts = dict(scheme=0, recon=0, flux=0, bp=1, rk=2, cfl=0.4, precision=0)
    a, b = wenofv.Batch(insts, **opts), wenofv.Batch(insts, **opts)
    b.set_schedule(persistent)
    for k, r in enumerate((8, 24, 64, 13)):
        b.set_rows_per_strip(k, r)
    for x in (a, b):
        for k, w in enumerate(wls):
            x.set_state(k, w.U0)
        x.advance(fixed_steps=3)
    for k in range(len(wls)):
        np.testing.assert_array_equal(a.get_state(k), b.get_state(k))


def test_autotune_leaves_recorded_limiter_flags_alone():
    """The tuner's timed stages write neither the limiter flags a context records nor its
    error bits: the flags read after tuning are those of the last stage of the run."""
    wl = cases.liquid_column(nx=160, ny=96, precision=0)
    sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    sol.set_state(wl.U0)
    L = wenofv.lib()
    assert L.wfv_record_limiter_flags(sol._h, 1) == wenofv.OK
    print("autotune keeps limiter flags", flush=True)
    sol.advance(fixed_steps=4)
    before = np.empty((wl.ny, wl.nx))
    assert L.wfv_get_limiter_flags(sol._h, before.ctypes.data) == wenofv.OK
    assert before.sum() > 0  # the shock front limits some cells
    sol.autotune(trials=1, cache=False)
    after = np.empty((wl.ny, wl.nx))
    assert L.wfv_get_limiter_flags(sol._h, after.ctypes.data) == wenofv.OK
    np.testing.assert_array_equal(after, before)
