"""Reflective walls (no reference counterpart: the reference's only boundary is periodic,
src/WENOFV.cpp:2249-2295) pinned to the periodic path, which the golden fixtures pin to the
reference.

A wall at a domain edge mirrors the three ghost layers and negates the normal momentum.  A
walled domain of n cells therefore sees exactly the values a periodic domain of 2n cells
holding the state and its mirror image sees: every stencil, face value, flux, CFL term and
RK update of the first n cells is the same arithmetic on the same operands.  The walled run
must equal the first half of the doubled periodic run bit for bit, in x and in y, for every
form of the scheme; and per-phase mass (the walls' normal mass flux vanishes) must be
conserved to rounding."""
import numpy as np
import pytest

import cases
import oracle as orc

PRM = dict(rho10=1.0, rho20=1e-2, sound1=10.0, sound2=10.0, preb=0.0)


def smooth_state(nx, ny, Lx, Ly, seed):
    """A smooth, non-symmetric two-phase state: a volume-fraction bump, a pressure wave and
    velocities of both signs (cell averages by 4x4 Gauss-Legendre)."""
    rng = np.random.default_rng(seed)
    cx, cy, ph = rng.uniform(0.3, 0.7) * Lx, rng.uniform(0.3, 0.7) * Ly, rng.uniform(0, 6)

    def pt(X, Y):
        r2 = ((X - cx) / Lx) ** 2 + ((Y - cy) / Ly) ** 2
        z1 = 0.2 + 0.6 * np.exp(-r2 / 0.03)
        p = 1.0 + 0.2 * np.sin(2 * np.pi * X / Lx + ph) * np.cos(np.pi * Y / Ly)
        u = 0.3 * np.sin(np.pi * Y / Ly + ph)
        v = -0.2 * np.cos(2 * np.pi * X / Lx)
        return cases.conservative(z1, p, u, v, **PRM)

    return cases.cell_average(nx, ny, (0.0, Lx, 0.0, Ly), pt)


def mirrored(U):
    """The periodic domain of twice the size in x and y holding U and its mirror images
    (normal momentum negated in each image)."""
    X = np.flip(U, axis=2).copy()
    X[2] *= -1.0
    top = np.concatenate([U, X], axis=2)
    Y = np.flip(top, axis=1).copy()
    Y[3] *= -1.0
    return np.concatenate([top, Y], axis=1)


FORMS = [(0, 0, 0, 0), (1, 0, 1, 1), (1, 1, 1, 1), (0, 2, 1, 1), (1, 3, 0, 1), (1, 1, 0, 0)]
# cell sizes 2^-4: every cell edge and width is exact in both domains
NX, NY, LX, LY = 20, 16, 1.25, 1.0


def pair(scheme, recon, flux, bp):
    opts = dict(scheme=scheme, recon=recon, flux=flux, bp=bp, rk=2, cfl=0.4, nranks=1, **PRM)
    walls = orc.Oracle(NX, NY, domain=(0.0, LX, 0.0, LY), bc=1, **opts)
    per = orc.Oracle(2 * NX, 2 * NY, domain=(0.0, 2 * LX, 0.0, 2 * LY), bc=0, **opts)
    return walls, per


@pytest.mark.parametrize("scheme,recon,flux,bp", FORMS)
def test_walls_rhs_equals_mirrored_periodic_domain(scheme, recon, flux, bp):
    """One right-hand side: bit for bit, face values and limiter included."""
    U = smooth_state(NX, NY, LX, LY, seed=3 + recon)
    walls, per = pair(scheme, recon, flux, bp)
    rw, fw, tw, ew = walls.assemble_rhs(U, want_faces=True)
    rp, fp, tp, ep = per.assemble_rhs(mirrored(U), want_faces=True)
    assert ew == ep == 0
    np.testing.assert_array_equal(rw, rp[:, :NY, :NX])
    np.testing.assert_array_equal(fw, fp[:NY, :NX])
    np.testing.assert_array_equal(tw, tp[:NY, :NX])
    assert walls.compute_dt(U) == per.compute_dt(mirrored(U))


@pytest.mark.parametrize("scheme,recon,flux,bp", FORMS)
def test_walls_time_loop_equals_mirrored_periodic_domain(scheme, recon, flux, bp):
    """Six RK3 steps: the image half sums its face fluxes in mirrored order, so the runs
    part at rounding level only."""
    U = smooth_state(NX, NY, LX, LY, seed=3 + recon)
    walls, per = pair(scheme, recon, flux, bp)
    Uw, dw, tw, ew = walls.run(U, fixed_steps=6)
    Up, dp, tp, ep = per.run(mirrored(U), fixed_steps=6)
    assert ew == ep == 0
    assert dw[0] == dp[0]
    np.testing.assert_allclose(dw, dp, rtol=1e-13)
    for v in range(4):
        np.testing.assert_allclose(Uw[v], Up[v, :NY, :NX], rtol=0, atol=1e-12 * np.abs(Up[v]).max())


def test_walls_conserve_phase_mass():
    nx, ny = 24, 20
    U = smooth_state(nx, ny, 1.0, 1.0, seed=9)
    o = orc.Oracle(nx, ny, domain=(0.0, 1.0, 0.0, 1.0), bc=1, scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4,
                   nranks=1, **PRM)
    U1, _, _, err = o.run(U, fixed_steps=20)
    assert err == 0
    m0, m1 = U[:2].sum(axis=(1, 2)), U1[:2].sum(axis=(1, 2))
    np.testing.assert_allclose(m1, m0, rtol=1e-13)


def mirrored_noslip(U):
    """As mirrored(), with no-slip images: both momenta negated once per mirror."""
    X = np.flip(U, axis=2).copy()
    X[2:] *= -1.0
    top = np.concatenate([U, X], axis=2)
    Y = np.flip(top, axis=1).copy()
    Y[2:] *= -1.0
    return np.concatenate([top, Y], axis=1)


@pytest.mark.parametrize("scheme,recon,flux,bp", FORMS + [(2, 0, 1, 1), (1, 4, 0, 1)])
def test_noslip_rhs_equals_mirrored_periodic_domain(scheme, recon, flux, bp):
    """No-slip walls (bc 2): the same construction with both momenta negated in the images."""
    U = smooth_state(NX, NY, LX, LY, seed=7 + recon)
    opts = dict(scheme=scheme, recon=recon, flux=flux, bp=bp, rk=2, cfl=0.4, nranks=1, **PRM)
    walls = orc.Oracle(NX, NY, domain=(0.0, LX, 0.0, LY), bc=2, **opts)
    per = orc.Oracle(2 * NX, 2 * NY, domain=(0.0, 2 * LX, 0.0, 2 * LY), bc=0, **opts)
    rw, fw, tw, ew = walls.assemble_rhs(U, want_faces=True)
    rp, fp, tp, ep = per.assemble_rhs(mirrored_noslip(U), want_faces=True)
    assert ew == ep == 0
    np.testing.assert_array_equal(rw, rp[:, :NY, :NX])
    np.testing.assert_array_equal(fw, fp[:NY, :NX])
    np.testing.assert_array_equal(tw, tp[:NY, :NX])


@pytest.mark.parametrize("scheme,recon,flux,bp", FORMS + [(2, 1, 1, 1), (1, 4, 1, 1)])
def test_fixed_ghost_rhs_equals_embedding_periodic_domain(scheme, recon, flux, bp):
    """Fixed ghost cells (bc 3, the reference's NEUMANN2): one RHS equals the interior of a
    larger periodic domain that holds the same values in the WFV_GHOST-deep ring (and
    anything further out, which no interior stencil reaches)."""
    G, E = orc.GHOST, 2 * orc.GHOST  # ghost layers; embedding margin
    big = smooth_state(NX + 2 * E, NY + 2 * E, LX * (NX + 2 * E) / NX, LY * (NY + 2 * E) / NY, seed=11 + recon)
    U = big[:, E:E + NY, E:E + NX]
    pad = big[:, E - G:E + NY + G, E - G:E + NX + G]
    opts = dict(scheme=scheme, recon=recon, flux=flux, bp=bp, rk=2, cfl=0.4, nranks=1, **PRM)
    fixed = orc.Oracle(NX, NY, domain=(0.0, LX, 0.0, LY), bc=3, ghost=pad, **opts)
    per = orc.Oracle(NX + 2 * E, NY + 2 * E, domain=(0.0, LX * (NX + 2 * E) / NX, 0.0, LY * (NY + 2 * E) / NY),
                     bc=0, **opts)
    rf, ff, tf, ef = fixed.assemble_rhs(U, want_faces=True)
    rp, fp, tp, ep = per.assemble_rhs(big, want_faces=True)
    assert ef == 0
    np.testing.assert_array_equal(rf, rp[:, E:E + NY, E:E + NX])
    np.testing.assert_array_equal(ff, fp[E:E + NY, E:E + NX])
    np.testing.assert_array_equal(tf, tp[E:E + NY, E:E + NX])
