"""Wavefront decomposition on the GPU: grids whose width leaves a narrow remainder strip,
which the stage kernels pack as several row segments per wavefront (strip_plan), against the
oracle.  All-FP64, tolerance 0: right-hand sides, face values and limiter theta of one
evaluation, then dt and state after a few RK3 steps, for widths from 1 column to a few
strips (the library's minimum is 3), several rows per strip and both stencil radii."""
import numpy as np
import pytest

import oracle as orc
import wenofv

pytestmark = pytest.mark.gpu

PRM = dict(rho10=1.0, rho20=0.1, sound1=10.0, sound2=10.0, preb=0.0)


def state(rng, nx, ny):
    X, Y = np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny)
    z1 = 0.5 + 0.3 * np.sin(2 * np.pi * X + 1.0) * np.cos(2 * np.pi * Y) + 0.02 * rng.standard_normal((ny, nx))
    p = 1.0 + 0.05 * rng.standard_normal((ny, nx))
    u = 0.3 + 0.1 * rng.standard_normal((ny, nx))
    v = -0.2 + 0.1 * rng.standard_normal((ny, nx))
    r1, r2 = PRM["rho10"] + p / PRM["sound1"] ** 2, PRM["rho20"] + p / PRM["sound2"] ** 2
    a, b = z1 * r1, (1 - z1) * r2
    return np.stack([a, b, (a + b) * u, (a + b) * v])


# (nx, ny, rows per strip): remainders of 3..57 columns packed 2..9 segments per wavefront,
# segments past the last row, a tail of idle lanes, and widths without a full strip
CASES = [(3, 9, 4), (5, 17, 3), (6, 20, 8), (20, 13, 64), (26, 31, 5), (27, 8, 1), (64, 45, 8), (64, 64, 64),
         (70, 37, 12), (117, 29, 2), (128, 50, 16), (140, 21, 7)]


@pytest.mark.parametrize("scheme,recon", [(1, 0), (0, 3), (2, 0)])
@pytest.mark.parametrize("nx,ny,ty", CASES, ids=lambda c: str(c))
def test_packed_remainder_strips_match_oracle(nx, ny, ty, scheme, recon):
    rng = np.random.default_rng(nx * 1000 + ny)
    U = state(rng, nx, ny)
    print(f"strips {nx}x{ny} ty{ty} s{scheme} r{recon}", flush=True)
    kw = dict(scheme=scheme, recon=recon, flux=1, bp=1, rk=2, cfl=0.4, domain=(0.0, 1.0, 0.0, 1.0), **PRM)
    # periodic, slip walls, no-slip walls; the oracle's walls mirror once, so they need
    # grids at least as wide as its 4-cell padding
    bc = (nx + ny) % 3 if min(nx, ny) >= 4 else 0
    sol = wenofv.Solver(nx, ny, bc=bc, geometry_ranks=1, rows_per_strip=ty, **kw)
    o = orc.Oracle(nx, ny, bc=bc, nranks=1, **kw)
    sol.set_state(U)
    rhs, faces, theta = sol.assemble_rhs(want_faces=True)
    r_o, f_o, t_o, err = o.assemble_rhs(U, want_faces=True)
    exact = scheme != 2  # WENO-EXP's device cube: 1e-12 (test_gpu_wenoexp.py)
    if exact:
        np.testing.assert_array_equal(rhs, r_o)
        np.testing.assert_array_equal(faces, f_o)
        np.testing.assert_array_equal(theta, t_o)
    else:
        np.testing.assert_allclose(rhs, r_o, rtol=0, atol=1e-12 * np.abs(r_o).max())
    n, t, dts = sol.advance(fixed_steps=3, max_dts=4)
    U_o, dts_o, t_o2, err = o.run(U, fixed_steps=3)
    assert err == 0 and n == 3 and sol.flags() == 0
    if exact:
        np.testing.assert_array_equal(dts, dts_o)
        np.testing.assert_array_equal(sol.get_state(), U_o)
    else:
        np.testing.assert_allclose(sol.get_state(), U_o, rtol=0, atol=1e-10 * np.abs(U_o).max())


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_of_narrow_grids_equals_single_solvers(precision):
    shapes = [(3, 9), (6, 20), (20, 13), (64, 64), (70, 37), (117, 29), (130, 66), (64, 512)]
    opts = dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=precision)
    rng = np.random.default_rng(3)
    insts, states = [], []
    for k, (nx, ny) in enumerate(shapes):
        insts.append(dict(nx=nx, ny=ny, bc=k % 3, domain=(0.0, 1.0, 0.0, 1.0), **PRM))
        states.append(state(rng, nx, ny))
    print(f"narrow batch p{precision}", flush=True)
    b = wenofv.Batch(insts, geometry_ranks=1, **opts)
    for k, U in enumerate(states):
        b.set_state(k, U)
    n, t, f = b.advance(fixed_steps=4)
    assert (n == 4).all() and (f == 0).all()
    for k, (nx, ny) in enumerate(shapes):
        sol = wenofv.Solver(nx, ny, bc=k % 3, geometry_ranks=1, domain=(0.0, 1.0, 0.0, 1.0), **PRM, **opts)
        sol.set_state(states[k])
        sol.advance(fixed_steps=4)
        np.testing.assert_array_equal(b.get_state(k), sol.get_state())
