"""Initial conditions of the benchmark workloads (host side, numpy).

Cell averages are 4x4 Gauss-Legendre means of the point state, as CWENOFV::initializeAve
does (src/WENOFV.cpp:232-279), with the point state built as TwoPhaseEquation::getU0
(include/Equations.hpp:979-1002): rho_k = (p - preb)/c_k^2 + rho_k0, U = (z1 rho1, z2 rho2,
rho u, rho v).

Workloads (the reference has none of them; choices are documented per function):
  tanh_interface   smooth tanh interface advected in a periodic unit box
  liquid_column    liquid column hit by a planar pressure step in gas, reflective walls
  perturbed_column the same column with a seeded 16-mode radius perturbation
  batch_instances  seeded independent liquid-column cases of uneven grid sizes
  multi_column     many seeded non-overlapping columns and a pressure step
"""
import numpy as np

# Gauss-Legendre-4 on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:64-75)
GL_PTS = np.array([-0.861136311594052575223946488893, -0.339981043584856264802665759103,
                   0.339981043584856264802665759103, 0.861136311594052575223946488893])
GL_WTS = np.array([0.347854845137453857373063949222, 0.652145154862546142626936050778,
                   0.652145154862546142626936050778, 0.347854845137453857373063949222])


class Workload:
    """A solver configuration plus its initial cell averages."""

    def __init__(self, name, nx, ny, domain, params, U0, bc, opts, steps=None, t_end=None):
        self.name, self.nx, self.ny, self.domain = name, nx, ny, domain
        self.params = params  # dict rho10, rho20, sound1, sound2, preb
        self.U0, self.bc, self.opts = U0, bc, opts
        self.steps, self.t_end = steps, t_end

    def solver_kwargs(self, **over):
        kw = dict(domain=self.domain, bc=self.bc, geometry_ranks=1, **self.params, **self.opts)
        kw.update(over)
        return kw


def cell_average(nx, ny, domain, point_state):
    """4x4 Gauss-Legendre cell averages of point_state(x, y) -> (4, ...) arrays."""
    xL, xR, yL, yR = domain
    dx, dy = (xR - xL) / nx, (yR - yL) / ny
    xc = xL + (np.arange(nx) + 0.5) * dx
    yc = yL + (np.arange(ny) + 0.5) * dy
    acc = np.zeros((4, ny, nx))
    for a in range(4):
        for b in range(4):
            X, Y = np.meshgrid(xc + 0.5 * dx * GL_PTS[a], yc + 0.5 * dy * GL_PTS[b])
            acc += point_state(X, Y) * (0.5 * GL_WTS[a] * 0.5 * GL_WTS[b])
    return acc


def conservative(z1, p, u, v, rho10, rho20, sound1, sound2, preb):
    r1 = (p - preb) / sound1 ** 2 + rho10
    r2 = (p - preb) / sound2 ** 2 + rho20
    a, b = r1 * z1, r2 * (1.0 - z1)
    return np.stack([a, b, (a + b) * u, (a + b) * v])


def shock_state(p0, ratio, rho20, sound2, preb=0.0):
    """Post-shock (p, u) of a right-moving pressure step into gas at rest.

    Rankine-Hugoniot for the linear gas EOS rho = rho20 + (p - preb)/c^2: the mass flux
    m satisfies m^2 = [p]/(1/rho_a - 1/rho_b) and u_b = m (1/rho_a - 1/rho_b)."""
    pa, pb = p0, p0 * ratio
    ra = rho20 + (pa - preb) / sound2 ** 2
    rb = rho20 + (pb - preb) / sound2 ** 2
    ub = np.sqrt((pb - pa) * (1.0 / ra - 1.0 / rb))
    return pb, ub


def tanh_interface(nx=128, ny=128):
    """Smooth interface advection in the periodic unit box: z1 = 0.5+0.5 tanh((0.2-r)/0.01)
    about (0.5, 0.5), clipped to [eps, 1-eps] with eps = 1e-8 (the reference's g_epsa of
    SMOOTH, include/Equations.hpp:505); rho10 = 1, rho20 = 0.1, c1 = c2 = 10 (the linear
    EOS's counterpart of "c0 ~ 10"), p = 1, (u, v) = (1, 0.5).  WENO-JS + LLF, RK3, CFL 0.4."""
    prm = dict(rho10=1.0, rho20=0.1, sound1=10.0, sound2=10.0, preb=0.0)
    eps = 1e-8

    def pt(X, Y):
        r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
        z1 = eps + (1 - 2 * eps) * (0.5 + 0.5 * np.tanh((0.2 - r) / 0.01))
        return conservative(z1, 1.0, 1.0, 0.5, **prm)

    U0 = cell_average(nx, ny, (0.0, 1.0, 0.0, 1.0), pt)
    return Workload("tanh_interface", nx, ny, (0.0, 1.0, 0.0, 1.0), prm, U0, 0,
                    dict(scheme=0, recon=0, flux=0, bp=0, rk=2, cfl=0.4), t_end=1.0)


def _column_state(prm, cx, cy, radius_fn, x_shock, ratio, z_in=1 - 1e-6, z_out=1e-6, p0=1.0):
    ps, us = shock_state(p0, ratio, prm["rho20"], prm["sound2"], prm["preb"])

    def pt(X, Y):
        r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
        th = np.arctan2(Y - cy, X - cx)
        z1 = np.where(r < radius_fn(th), z_in, z_out)
        post = X < x_shock
        p = np.where(post, ps, p0)
        u = np.where(post, us, 0.0)
        return conservative(z1, p, u, 0.0 * X, **prm)

    return pt


def liquid_column(nx=1024, ny=512, precision=0):
    """Liquid column (radius 0.15 at (0.5, 0.5), z1 = 1-1e-6, rho10 = 1) in gas (z1 = 1e-6,
    rho20 = 1e-3) on [0,2]x[0,1] with reflective walls; c1 = c2 = 10; pre-shock p = 1, u = 0;
    a pressure step of ratio 2 at x = 0.2 with the Rankine-Hugoniot velocity of the gas.
    WENO-Z + HLLC + bound-preserving limiter, RK3, CFL 0.4, 2000 steps."""
    prm = dict(rho10=1.0, rho20=1e-3, sound1=10.0, sound2=10.0, preb=0.0)
    dom = (0.0, 2.0, 0.0, 1.0)
    U0 = cell_average(nx, ny, dom, _column_state(prm, 0.5, 0.5, lambda th: 0.15, 0.2, 2.0))
    return Workload("liquid_column", nx, ny, dom, prm, U0, 1,
                    dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=precision), steps=2000)


def perturbed_column(nx=4096, ny=2048, precision=1, seed=3):
    """liquid_column geometry with radius 0.15 (1 + sum_{m=1..16} a_m sin(m th + phi_m)),
    a_m and phi_m from `seed`, scaled so that max |sum| = 0.01.  200 steps."""
    prm = dict(rho10=1.0, rho20=1e-3, sound1=10.0, sound2=10.0, preb=0.0)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, 16)
    ph = rng.uniform(0, 2 * np.pi, 16)
    th = np.linspace(0, 2 * np.pi, 4096, endpoint=False)
    s = sum(a[m] * np.sin((m + 1) * th + ph[m]) for m in range(16))
    a *= 0.01 / np.abs(s).max()

    # radius on a fine periodic angle table (linear interpolation error < 1e-9)
    tab_t = np.linspace(-np.pi, np.pi, 1 << 16)
    tab_r = 0.15 * (1 + sum(a[m] * np.sin((m + 1) * tab_t + ph[m]) for m in range(16)))

    def radius(t):
        return np.interp(t, tab_t, tab_r)

    dom = (0.0, 2.0, 0.0, 1.0)
    U0 = cell_average(nx, ny, dom, _column_state(prm, 0.5, 0.5, radius, 0.2, 2.0))
    return Workload("perturbed_column", nx, ny, dom, prm, U0, 1,
                    dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=precision), steps=200)


def batch_specs(n=256, seed=4):
    """Seeded instance list: nx, ny independently from {64,128,256,512}; column radius
    U(0.05,0.25), centre U(0.3,0.7)^2; pressure ratio log-uniform in [1.2, 5]; rho20/rho10 in
    {1e-3, 1e-2, 0.1}.  Unit box with reflective walls; the step sits halfway between the
    left wall and the column."""
    rng = np.random.default_rng(seed)
    sizes = np.array([64, 128, 256, 512])
    out = []
    for _ in range(n):
        nx, ny = int(rng.choice(sizes)), int(rng.choice(sizes))
        r = rng.uniform(0.05, 0.25)
        cx, cy = rng.uniform(0.3, 0.7, 2)
        ratio = float(np.exp(rng.uniform(np.log(1.2), np.log(5.0))))
        dr = float(rng.choice([1e-3, 1e-2, 0.1]))
        out.append(dict(nx=nx, ny=ny, radius=r, cx=cx, cy=cy, ratio=ratio, rho20=dr, x_shock=0.5 * (cx - r)))
    return out


def batch_instance(spec, precision=1):
    prm = dict(rho10=1.0, rho20=spec["rho20"], sound1=10.0, sound2=10.0, preb=0.0)
    dom = (0.0, 1.0, 0.0, 1.0)
    U0 = cell_average(spec["nx"], spec["ny"], dom,
                      _column_state(prm, spec["cx"], spec["cy"], lambda th: spec["radius"], spec["x_shock"],
                                    spec["ratio"]))
    return Workload("batch_instance", spec["nx"], spec["ny"], dom, prm, U0, 1,
                    dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=precision), steps=500)


def multi_column(nx=8192, ny=8192, ncol=64, seed=5, precision=1):
    """64 columns of radius U(0.01, 0.04) at seeded non-overlapping positions (rejection
    sampling in [0.15, 0.95] x [0.05, 0.95]) in the walled unit box; pressure-ratio-3 step
    at x = 0.1.  200 steps."""
    prm = dict(rho10=1.0, rho20=1e-3, sound1=10.0, sound2=10.0, preb=0.0)
    rng = np.random.default_rng(seed)
    cols = []
    while len(cols) < ncol:
        r = rng.uniform(0.01, 0.04)
        c = rng.uniform([0.15 + r, 0.05 + r], [0.95 - r, 0.95 - r])
        if all(np.hypot(*(c - q[:2])) > r + q[2] + 0.005 for q in cols):
            cols.append(np.array([c[0], c[1], r]))
    cols = np.array(cols)
    ps, us = shock_state(1.0, 3.0, prm["rho20"], prm["sound2"])
    x_shock = 0.1

    def pt(X, Y, subset=cols):
        inside = np.zeros(X.shape, dtype=bool)
        for cx, cy, r in subset:
            inside |= (X - cx) ** 2 + (Y - cy) ** 2 < r * r
        z1 = np.where(inside, 1 - 1e-6, 1e-6)
        post = X < x_shock
        return conservative(z1, np.where(post, ps, 1.0), np.where(post, us, 0.0), 0.0 * X, **prm)

    dom = (0.0, 1.0, 0.0, 1.0)
    # The state is uniform away from the columns and the step: fill the two uniform states,
    # then average only the cell blocks that the columns or the step touch (the same
    # quadrature as cell_average on those cells).
    dx, dy = 1.0 / nx, 1.0 / ny
    xc = (np.arange(nx) + 0.5) * dx
    U0 = np.empty((4, ny, nx))
    U0[:] = conservative(1e-6, 1.0, 0.0, 0.0, **prm)[:, None, None]
    U0[:, :, xc + 0.5 * dx <= x_shock] = conservative(1e-6, ps, us, 0.0, **prm)[:, None, None]

    def average_block(i0, i1, j0, j1):
        i0, j0, i1, j1 = max(i0, 0), max(j0, 0), min(i1, nx), min(j1, ny)
        xs = (np.arange(i0, i1) + 0.5) * dx
        ys = (np.arange(j0, j1) + 0.5) * dy
        near = [q for q in cols if q[0] + q[2] >= i0 * dx and q[0] - q[2] <= i1 * dx
                and q[1] + q[2] >= j0 * dy and q[1] - q[2] <= j1 * dy]
        acc = np.zeros((4, j1 - j0, i1 - i0))
        for a in range(4):
            for b in range(4):
                X, Y = np.meshgrid(xs + 0.5 * dx * GL_PTS[a], ys + 0.5 * dy * GL_PTS[b])
                acc += pt(X, Y, near) * (0.5 * GL_WTS[a] * 0.5 * GL_WTS[b])
        U0[:, j0:j1, i0:i1] = acc

    ishock = int(x_shock / dx)
    for j in range(0, ny, 1024):
        average_block(ishock - 1, ishock + 2, j, j + 1024)
    for cx, cy, r in cols:
        average_block(int((cx - r) / dx) - 1, int((cx + r) / dx) + 2, int((cy - r) / dy) - 1, int((cy + r) / dy) + 2)
    return Workload("multi_column", nx, ny, dom, prm, U0, 1,
                    dict(scheme=1, recon=0, flux=1, bp=1, rk=2, cfl=0.4, precision=precision), steps=200)
