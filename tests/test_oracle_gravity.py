"""The gravity source term (assembleSourceTerm, src/WENOFV.cpp:1831-1872: rhs[3] -= rho g with
g = 9.81 for DAMBREAK and RTI, 2 pi for STANDING_WAVE) in the oracle.

Pinned two ways:
- against the reference's own solver: tests/golden/solver_gravity.npz holds right-hand sides
  (2 option sets) and 20-step runs of the three gravity cases on a periodic 4-rank grid,
  written by the unmodified reference with its source term switched on (make_golden.py,
  `gravity`); the oracle reproduces every one bit for bit;
- by the closed form: a uniform state at rest, in a slip-walled or periodic box whose cell
  widths are exact (1/16), has rhs = (0, 0, 0, -rho g) exactly, because every face carries the
  same flux, which cancels between the two sides of each cell."""
import numpy as np
import pytest

import oracle as orc
from conftest import golden

G = golden("solver_gravity.npz")
GRAVITY = {"dambreak": 9.81, "rti": 9.81, "standing_wave": 2 * np.pi}


def oracle_for(U, g, **kw):
    p = G["params"]
    ny, nx = U.shape[1:]
    return orc.Oracle(nx, ny, bc=0, rho10=p[0], rho20=p[1], sound1=p[2], sound2=p[3], preb=p[4], gravity=g, **kw)


def rhs_tags():
    return sorted(k for k in G.files if k.startswith("rhs_") and k != "rhs_U")


@pytest.mark.parametrize("tag", rhs_tags())
def test_gravity_rhs_matches_reference(tag):
    name = tag[len("rhs_"):].rsplit("_s", 1)[0]
    s, r, f, b = tag.rsplit("_", 4)[1:]
    o = oracle_for(G["rhs_U"], GRAVITY[name], scheme=int(s[1]), recon=int(r[1]), flux=int(f[1]), bp=int(b[1]))
    rhs, _, _, err = o.assemble_rhs(G["rhs_U"])
    assert err == 0
    np.testing.assert_array_equal(rhs, G[tag])
    # the source term is there: without it only the vertical momentum changes
    rhs0, _, _, _ = oracle_for(G["rhs_U"], 0.0, scheme=int(s[1]), recon=int(r[1]), flux=int(f[1]),
                               bp=int(b[1])).assemble_rhs(G["rhs_U"])
    np.testing.assert_array_equal(rhs0[:3], rhs[:3])
    assert (rhs0[3] > rhs[3]).all()


@pytest.mark.parametrize("name", sorted(GRAVITY))
def test_gravity_run_matches_reference(name):
    o = oracle_for(G["run_U"], GRAVITY[name], scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4)
    U, dts, t, err = o.run(G["run_U"], fixed_steps=20)
    assert err == 0
    np.testing.assert_array_equal(dts, G[f"run_{name}_dts"])
    np.testing.assert_array_equal(U, G[f"run_{name}_final"])


def uniform_at_rest(n, z1=0.3, p=1.5, prm=(1.2, 1.0, 1.0, 1.0, 0.0)):
    rho10, rho20, c1, c2, preb = prm
    r1, r2 = rho10 + (p - preb) / c1 ** 2, rho20 + (p - preb) / c2 ** 2
    U = np.zeros((4, n, n))
    U[0], U[1] = z1 * r1, (1 - z1) * r2
    return U


@pytest.mark.parametrize("bc", [0, 1, 2])
@pytest.mark.parametrize("scheme,recon,flux,bp", [(1, 0, 1, 1), (0, 1, 0, 0), (1, 3, 1, 1), (2, 4, 1, 1)])
@pytest.mark.parametrize("g", [9.81, 2 * np.pi])
def test_uniform_state_at_rest_feels_only_gravity(bc, scheme, recon, flux, bp, g):
    U = uniform_at_rest(16)
    o = oracle_for(U, g, scheme=scheme, recon=recon, flux=flux, bp=bp, nranks=1)
    o.p.bc = bc
    rhs, _, _, err = o.assemble_rhs(U)
    assert err == 0
    np.testing.assert_array_equal(rhs[:3], 0.0)
    np.testing.assert_array_equal(rhs[3], -(U[0] + U[1]) * g)
