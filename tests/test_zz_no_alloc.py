"""The allocation-free entry point wfv_advance_fixed_noalloc, called through the C ABI on a
caller-owned device state, against wfv_advance on a context with the same options: bitwise
equal states, step counts and final times, on two calls in a row (the second on a smaller
grid, so it runs on scratch memory the first call left behind), in both precision policies
and for SSP-RK1 and SSP-RK3.  Also: the refusals of a grid too large for the scratch, of the
fixed-ghost boundary and of a non-positive step count."""
import ctypes

import numpy as np
import pytest

import wenofv

pytestmark = pytest.mark.gpu


def column_state(nx, ny, seed):
    """A liquid column in gas with a seeded velocity field, [4][ny][nx]."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    r = np.hypot(X - 0.45, Y - 0.55)
    z1 = 1e-6 + (1 - 2e-6) * 0.5 * (1 + np.tanh((0.25 - r) / 0.03))
    rho1, rho2 = 1.0 + 0.01 * rng.random(), 0.1
    a, b = z1 * rho1, (1 - z1) * rho2
    rho = a + b
    u = 0.3 * np.sin(2 * np.pi * Y) + 0.05 * rng.standard_normal((ny, nx))
    v = -0.2 * np.cos(2 * np.pi * X)
    return np.ascontiguousarray(np.stack([a, b, rho * u, rho * v]))


KW = dict(domain=(0.0, 1.0, 0.0, 1.0), rho10=1.0, rho20=0.1, sound1=10.0, sound2=10.0, scheme=1, recon=0, flux=1,
          bp=1, cfl=0.4, bc=1)


def hip():
    """The HIP runtime libwenofv.so is linked against (already loaded with it): the caller's
    device buffers come from it, as an application's would."""
    wenofv.lib()
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipFree.argtypes = [ctypes.c_void_p]
    return L


H2D, D2H = 1, 2


def noalloc(sol, U0, steps):
    H = hip()
    U0 = np.ascontiguousarray(U0, dtype=np.float64)
    dU = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dU), U0.nbytes) == 0
    try:
        assert H.hipMemcpy(dU, U0.ctypes.data, U0.nbytes, H2D) == 0
        n = ctypes.c_int()
        t = ctypes.c_double()
        rc = wenofv.lib().wfv_advance_fixed_noalloc(ctypes.byref(sol.opt), ctypes.byref(sol.case), dU, 0.0, steps,
                                                    None, ctypes.byref(n), ctypes.byref(t))
        assert rc == wenofv.OK, wenofv.lib().wfv_last_error(None).decode()
        U = np.empty_like(U0)
        assert H.hipMemcpy(U.ctypes.data, dU, U0.nbytes, D2H) == 0
    finally:
        H.hipFree(dU)
    return U, n.value, t.value


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("rk", [0, 2])
def test_noalloc_matches_advance_on_two_calls(precision, rk):
    for nx, ny, steps, seed in ((152, 96, 12, 1), (40, 24, 7, 2)):
        print(f"noalloc vs advance: {nx}x{ny} precision={precision} rk={rk} steps={steps}", flush=True)
        U0 = column_state(nx, ny, seed)
        sol = wenofv.Solver(nx, ny, rk=rk, precision=precision, **KW)
        sol.set_state(U0)
        n_ref, t_ref, _ = sol.advance(fixed_steps=steps)
        U_ref = sol.get_state()
        U, n, t = noalloc(sol, U0, steps)
        assert n == n_ref == steps
        assert t == t_ref
        np.testing.assert_array_equal(U, U_ref)
        assert np.isfinite(U).all() and not np.array_equal(U, U0)
        sol.close()


def test_noalloc_refusals():
    L = wenofv.lib()
    # every refusal happens on the host, before any device work: the state pointer is
    # never dereferenced
    ptr = ctypes.c_void_p(0x1000)
    # 2048 x 1024 needs 128 MiB of registers: refused before any device work
    big = wenofv.Options(2048, 1024, 2, 0.4, 1, 1, 1, 0, 1, 1, 1, 0)
    case = wenofv.Case(0.0, 1.0, 0.0, 1.0, 1.0, 0.1, 10.0, 10.0, 0.0, 0.0)
    assert L.wfv_advance_fixed_noalloc(ctypes.byref(big), ctypes.byref(case), ptr, 0.0, 1, None, None, None) \
        == wenofv.ERR_ARG
    assert b"too large" in L.wfv_last_error(None)
    ghost = wenofv.Options(32, 32, 2, 0.4, 1, 1, 1, 0, 1, 3, 1, 0)
    assert L.wfv_advance_fixed_noalloc(ctypes.byref(ghost), ctypes.byref(case), ptr, 0.0, 1, None, None, None) \
        == wenofv.ERR_CONFIG
    ok = wenofv.Options(32, 32, 2, 0.4, 1, 1, 1, 0, 1, 1, 1, 0)
    assert L.wfv_advance_fixed_noalloc(ctypes.byref(ok), ctypes.byref(case), ptr, 0.0, 0, None, None, None) \
        == wenofv.ERR_ARG
