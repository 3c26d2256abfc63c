"""Mixed-precision solver against the all-FP64 GPU solver (itself bit-equal to the oracle).

Tolerance (stated by the workload definition): relative L1 <= 1e-5 per conserved field and
for z1, rho, u, v after the run; z1 stays in [0, 1] exactly; per-phase mass drift reported."""
import numpy as np
import pytest

import cases
import wenofv

pytestmark = pytest.mark.gpu

TOL = 1e-5


def derived(U, rho10, rho20, sound1, sound2, **_):
    c1, c2 = sound1 ** 2, sound2 ** 2
    q = rho20 * c2 - rho10 * c1
    qs = U[1] * c2 - U[0] * c1
    r = ((q - qs) + np.sqrt((q - qs) ** 2 + 4.0 * U[0] * c1 * U[1] * c2)) / (2.0 * U[1] * c2)
    z1 = r / (1 + r)
    rho = U[0] + U[1]
    return dict(z1=z1, rho=rho, u=U[2] / rho, v=U[3] / rho)


def rel_l1(a, b, scale=None):
    """sum|a-b| / sum|scale| (scale defaults to b)."""
    return np.abs(a - b).sum() / max(np.abs(b if scale is None else scale).sum(), 1e-300)


def field_errors(U32, U64, params):
    """Relative L1 errors of the four conserved fields and of z1, rho, u, v.  Momentum and
    velocity components are normalised by the magnitude of their vector (a component that
    vanishes in the FP64 run, e.g. by symmetry, has no scale of its own)."""
    d32, d64 = derived(U32, **params), derived(U64, **params)
    mom = np.hypot(U64[2], U64[3])
    vel = np.hypot(d64["u"], d64["v"])
    return {"z1rho1": rel_l1(U32[0], U64[0]), "z2rho2": rel_l1(U32[1], U64[1]),
            "rhou": rel_l1(U32[2], U64[2], mom), "rhov": rel_l1(U32[3], U64[3], mom),
            "z1": rel_l1(d32["z1"], d64["z1"]), "rho": rel_l1(d32["rho"], d64["rho"]),
            "u": rel_l1(d32["u"], d64["u"], vel), "v": rel_l1(d32["v"], d64["v"], vel)}


@pytest.mark.parametrize("recon", [0, 1, 2, 3])
@pytest.mark.parametrize("scheme", [0, 1])
def test_mixed_matches_fp64_liquid_column(recon, scheme):
    wl = cases.liquid_column(nx=256, ny=128)
    print(f"mixed vs fp64 recon{recon} scheme{scheme}", flush=True)
    out = {}
    for prec in (0, 1):
        sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs(precision=prec, recon=recon, scheme=scheme))
        sol.set_state(wl.U0)
        n, t, dts = sol.advance(fixed_steps=60, max_dts=64)
        out[prec] = (sol.get_state(), dts, sol.flags())
    U64, U32 = out[0][0], out[1][0]
    assert out[1][2] == 0 and np.isfinite(U32).all()
    err = field_errors(U32, U64, wl.params)
    print({k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert max(err.values()) <= TOL, err
    d32 = derived(U32, **wl.params)
    assert d32["z1"].min() >= 0.0 and d32["z1"].max() <= 1.0
    m0 = wl.U0[:2].sum(axis=(1, 2))
    np.testing.assert_allclose(U32[:2].sum(axis=(1, 2)), m0, rtol=1e-12)


def test_mixed_tanh_interface_periodic():
    wl = cases.tanh_interface(64, 64)
    res = {}
    for prec in (0, 1):
        sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs(precision=prec))
        sol.set_state(wl.U0)
        sol.advance(t_end=0.05)
        res[prec] = sol.get_state()
    err = field_errors(res[1], res[0], wl.params)
    print({k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert max(err.values()) <= TOL, err
