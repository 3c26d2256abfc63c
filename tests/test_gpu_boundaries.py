"""Boundary kinds on the GPU: no-slip walls (2) and fixed ghost cells (3, the reference's
NEUMANN2) against the oracle (whose walls are pinned to the reference's periodic path by
tests/test_oracle_walls.py), batches with every kind against single solvers, and the
reference's non-periodic test cases (DAMBREAK, DAMBREAK2, DAMBREAK3, RTI, STANDING_WAVE with
gravity; ZALESAK with ghost cells from its initial state) stepped on the GPU and on the
oracle.  Tolerance 0 for the 5-point forms in FP64.  Primitive3's THINC kernel uses the
device's exp, log and tanh, whose last bits may differ from glibc's; the divergence scales
such face-value differences by 1/(dx dy) (~1300 on these grids), so its right-hand sides and
states are compared to 1e-10 relative to each field's largest magnitude."""
import numpy as np
import pytest

import oracle as orc
import wenofv

pytestmark = pytest.mark.gpu

PRM = dict(rho10=1.0, rho20=1e-2, sound1=10.0, sound2=10.0, preb=0.0)


def smooth(nx, ny, Lx, Ly, seed):
    rng = np.random.default_rng(seed)
    cx, cy, ph = rng.uniform(0.3, 0.7) * Lx, rng.uniform(0.3, 0.7) * Ly, rng.uniform(0, 6)
    X, Y = np.meshgrid((np.arange(nx) + 0.5) * Lx / nx, (np.arange(ny) + 0.5) * Ly / ny)
    z1 = 0.2 + 0.6 * np.exp(-(((X - cx) / Lx) ** 2 + ((Y - cy) / Ly) ** 2) / 0.03)
    p = 1.0 + 0.2 * np.sin(2 * np.pi * X / Lx + ph) * np.cos(np.pi * Y / Ly)
    u = 0.3 * np.sin(np.pi * Y / Ly + ph)
    v = -0.2 * np.cos(2 * np.pi * X / Lx)
    r1 = PRM["rho10"] + (p - PRM["preb"]) / PRM["sound1"] ** 2
    r2 = PRM["rho20"] + (p - PRM["preb"]) / PRM["sound2"] ** 2
    a, b = z1 * r1, (1 - z1) * r2
    return np.stack([a, b, (a + b) * u, (a + b) * v])


def padded(nx, ny, Lx, Ly, seed):
    """A smooth state on the grid padded by GHOST cells (same spacing)."""
    G = wenofv.GHOST
    dx, dy = Lx / nx, Ly / ny
    return smooth(nx + 2 * G, ny + 2 * G, dx * (nx + 2 * G), dy * (ny + 2 * G), seed)


def compare(got, ref, exact):
    if exact:
        np.testing.assert_array_equal(got, ref)
    else:
        for k in range(ref.shape[0]):
            assert np.abs(got[k] - ref[k]).max() <= 1e-10 * max(np.abs(ref[k]).max(), 1e-300)


FORMS = [(0, 0, 0, 0), (1, 0, 1, 1), (1, 1, 1, 1), (0, 3, 1, 1), (2, 0, 1, 1), (1, 4, 1, 1)]


@pytest.mark.parametrize("bc", [2, 3])
@pytest.mark.parametrize("form", FORMS)
def test_rhs_and_steps_match_oracle(form, bc):
    scheme, recon, flux, bp = form
    nx, ny, Lx, Ly = 70, 37, 2.0, 1.0
    print(f"boundary {bc} form {form}", flush=True)
    pad = padded(nx, ny, Lx, Ly, seed=3 + recon)
    G = wenofv.GHOST
    U = pad[:, G:G + ny, G:G + nx].copy()
    kw = dict(scheme=scheme, recon=recon, flux=flux, bp=bp, rk=2, cfl=0.4, domain=(0.0, Lx, 0.0, Ly), **PRM)
    o = orc.Oracle(nx, ny, bc=bc, nranks=1, ghost=pad if bc == 3 else None, **kw)
    sol = wenofv.Solver(nx, ny, bc=bc, geometry_ranks=1, rows_per_strip=8, **kw)
    if bc == 3:
        sol.set_ghost_cells(pad)
    sol.set_state(U)
    exact = recon != 4
    rhs, faces, theta = sol.assemble_rhs(want_faces=True)
    r_o, f_o, t_o, err = o.assemble_rhs(U, want_faces=True)
    assert err == 0
    compare(rhs, r_o, exact)
    compare(faces.reshape(ny, nx, 4, 4, 4).transpose(3, 0, 1, 2, 4), f_o.reshape(ny, nx, 4, 4, 4).transpose(3, 0, 1, 2, 4),
            exact)
    n, t, dts = sol.advance(fixed_steps=4, max_dts=8)
    U_o, dts_o, t_o2, err = o.run(U, fixed_steps=4)
    assert err == 0 and sol.flags() == 0
    if exact:
        np.testing.assert_array_equal(dts, dts_o)
    compare(sol.get_state(), U_o, exact)


def test_batch_with_every_boundary_equals_single_solvers():
    shapes = [(64, 40), (90, 33), (57, 70), (80, 64)]
    opts = dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=0)
    insts, states, pads = [], [], []
    for k, (nx, ny) in enumerate(shapes):
        insts.append(dict(nx=nx, ny=ny, bc=k, domain=(0.0, 1.0, 0.0, 1.0), **PRM))
        pad = padded(nx, ny, 1.0, 1.0, seed=20 + k)
        pads.append(pad)
        states.append(pad[:, 4:4 + ny, 4:4 + nx].copy())
    b = wenofv.Batch(insts, geometry_ranks=1, **opts)
    for k in range(len(shapes)):
        b.set_state(k, states[k])
    b.set_ghost_cells(3, pads[3])
    n, t, f = b.advance(fixed_steps=6)
    assert (n == 6).all() and (f == 0).all()
    for k, (nx, ny) in enumerate(shapes):
        sol = wenofv.Solver(nx, ny, bc=k, geometry_ranks=1, domain=(0.0, 1.0, 0.0, 1.0), **PRM, **opts)
        if k == 3:
            sol.set_ghost_cells(pads[3])
        sol.set_state(states[k])
        sol.advance(fixed_steps=6)
        np.testing.assert_array_equal(b.get_state(k), sol.get_state())


def test_ghost_cells_need_boundary_3():
    sol = wenofv.Solver(16, 16, bc=1, **PRM)
    with pytest.raises(wenofv.WenofvError):
        sol.set_ghost_cells(np.zeros((4, 24, 24)))


@pytest.mark.parametrize("testcase", [3, 4, 5, 6, 7, 8])
def test_reference_cases_with_gravity_and_ghost_cells(testcase):
    """The reference's non-periodic cases (its MPI solver cannot run them): the GPU follows
    the oracle bit for bit for 30 steps, z1 stays in [0, 1] and, behind walls, each phase's
    mass is conserved."""
    kw, t_end = wenofv.reference_case(testcase)
    bc = kw.pop("bc")
    nx, ny = (64, 64) if testcase != 6 else (32, 128)
    U0 = wenofv.reference_initial_state(testcase, nx, ny, geometry_ranks=4)
    pad = wenofv.reference_initial_state_padded(testcase, nx, ny, geometry_ranks=4) if bc == 3 else None
    opts = dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4)
    print(f"reference case {testcase} bc {bc} {nx}x{ny}", flush=True)
    sol = wenofv.Solver(nx, ny, bc=bc, geometry_ranks=4, **kw, **opts)
    if pad is not None:
        sol.set_ghost_cells(pad)
    sol.set_state(U0)
    n, t, dts = sol.advance(fixed_steps=30, max_dts=32)
    U = sol.get_state()
    o = orc.Oracle(nx, ny, bc=bc, nranks=4, ghost=pad, **kw, **opts)
    U_o, dts_o, t_o, err = o.run(U0, fixed_steps=30)
    assert err == 0 and sol.flags() == 0
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(U, U_o)
    assert (U[0] >= 0).all() and (U[1] >= 0).all()  # partial densities stay non-negative
    if bc in (1, 2):
        m0, m1 = U0[:2].sum(axis=(1, 2)), U[:2].sum(axis=(1, 2))
        np.testing.assert_allclose(m1, m0, rtol=1e-12)
