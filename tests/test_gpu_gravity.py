"""The gravity source term on the GPU (assembleSourceTerm, src/WENOFV.cpp:1831-1872), through
the C ABI: the all-FP64 solver reproduces the reference's own right-hand sides and 20-step
runs of DAMBREAK, RTI and STANDING_WAVE gravity (tests/golden/solver_gravity.npz, periodic
4-rank grid) bit for bit; and in both precisions a uniform state at rest in a box with exact
cell widths has rhs = (0, 0, 0, -rho g) exactly, periodic or walled."""
import numpy as np
import pytest

import wenofv
from conftest import golden
from test_oracle_gravity import GRAVITY, rhs_tags, uniform_at_rest

pytestmark = pytest.mark.gpu

G = golden("solver_gravity.npz")


def solver_for(U, g, **kw):
    p = G["params"]
    ny, nx = U.shape[1:]
    return wenofv.Solver(nx, ny, rho10=p[0], rho20=p[1], sound1=p[2], sound2=p[3], preb=p[4], gravity=g, **kw)


@pytest.mark.parametrize("tag", rhs_tags())
def test_gravity_rhs_matches_reference_bitwise(tag):
    name = tag[len("rhs_"):].rsplit("_s", 1)[0]
    s, r, f, b = tag.rsplit("_", 4)[1:]
    print(f"gravity rhs {tag}", flush=True)
    sol = solver_for(G["rhs_U"], GRAVITY[name], scheme=int(s[1]), recon=int(r[1]), flux=int(f[1]), bp=int(b[1]))
    sol.set_state(G["rhs_U"])
    rhs, _, _ = sol.assemble_rhs()
    np.testing.assert_array_equal(rhs, G[tag])


@pytest.mark.parametrize("name", sorted(GRAVITY))
def test_gravity_run_matches_reference_bitwise(name):
    print(f"gravity run {name}", flush=True)
    sol = solver_for(G["run_U"], GRAVITY[name], scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4)
    sol.set_state(G["run_U"])
    n, t, dts = sol.advance(fixed_steps=20, max_dts=20)
    assert n == 20 and sol.flags() == 0
    np.testing.assert_array_equal(dts, G[f"run_{name}_dts"])
    np.testing.assert_array_equal(sol.get_state(), G[f"run_{name}_final"])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("bc", [0, 1, 2])
@pytest.mark.parametrize("scheme,recon,flux,bp", [(1, 0, 1, 1), (0, 1, 0, 0), (1, 3, 1, 1), (2, 4, 1, 1)])
def test_uniform_state_at_rest_feels_only_gravity(precision, bc, scheme, recon, flux, bp):
    U = uniform_at_rest(16)
    for g in (9.81, 2 * np.pi):
        sol = solver_for(U, g, scheme=scheme, recon=recon, flux=flux, bp=bp, bc=bc, precision=precision,
                         geometry_ranks=1)
        sol.set_state(U)
        rhs, _, _ = sol.assemble_rhs()
        np.testing.assert_array_equal(rhs[:3], 0.0)
        np.testing.assert_array_equal(rhs[3], -(U[0] + U[1]) * g)
