"""ctypes binding of the CPU FP64 restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline
leg of bench.py, always as the checker.  The product library never loads it.

State arrays are float64 numpy arrays of shape (4, ny, nx) (SoA, x fastest), holding
(z1*rho1, z2*rho2, rho*u, rho*v).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

ERR_NAN_FLUX, ERR_BAD_THETA, ERR_NAN_DT = 1, 2, 4

_dp = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("scheme", ctypes.c_int),
        ("recon", ctypes.c_int), ("flux", ctypes.c_int), ("bp", ctypes.c_int),
        ("bc", ctypes.c_int), ("rk", ctypes.c_int), ("cfl", ctypes.c_double),
        ("dX", ctypes.c_double), ("dY", ctypes.c_double),
        ("rho10", ctypes.c_double), ("rho20", ctypes.c_double),
        ("sound1", ctypes.c_double), ("sound2", ctypes.c_double), ("preb", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("dxC", _dp), ("dxW", _dp), ("dyC", _dp), ("ghost", _dp),
    ]

GHOST = 4  # ghost layers of a fixed-ghost boundary (bc 3)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    P = ctypes.POINTER(Params)
    lib.orc_weno.restype = ctypes.c_double
    lib.orc_weno.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
    lib.orc_weno7.restype = ctypes.c_double
    lib.orc_weno7.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double]
    for name in ("orc_weno_exp", "orc_weno_exp_thinc"):
        getattr(lib, name).restype = ctypes.c_double
        getattr(lib, name).argtypes = [_dp, ctypes.c_int, ctypes.c_double]
    for name in ("orc_volume_frac", "orc_pressure", "orc_sound_speed"):
        getattr(lib, name).restype = ctypes.c_double
        getattr(lib, name).argtypes = [P, _dp]
    lib.orc_max_eig.restype = ctypes.c_double
    lib.orc_max_eig.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double]
    lib.orc_phy_flux.argtypes = [P, _dp, _dp, ctypes.c_double, ctypes.c_double]
    lib.orc_riemann.argtypes = [P, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.orc_eigen_L.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_eigen_R.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_bp_theta.restype = ctypes.c_double
    lib.orc_bp_theta.argtypes = [P, _dp, _dp]
    lib.orc_geometry.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int] + [_dp] * 5
    lib.orc_assemble_rhs.restype = ctypes.c_int
    lib.orc_assemble_rhs.argtypes = [P, _dp, _dp, _dp, _dp]
    lib.orc_compute_dt.restype = ctypes.c_int
    lib.orc_compute_dt.argtypes = [P, _dp, _dp]
    lib.orc_rk_step.restype = ctypes.c_int
    lib.orc_rk_step.argtypes = [P, _dp, ctypes.c_double]
    lib.orc_run.restype = ctypes.c_int
    lib.orc_run.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int,
                            ctypes.POINTER(ctypes.c_int), _dp]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _vec(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if n is not None:
        assert a.size == n
    return a


def geometry(nx, ny, xL, xR, yL, yR, nranks=4):
    """Per-column/row widths and centres of the reference's 4-rank slab grid."""
    dxC, dxW, xc = (np.zeros(nx) for _ in range(3))
    dyC, yc = np.zeros(ny), np.zeros(ny)
    lib().orc_geometry(nx, ny, xL, xR, yL, yR, nranks, _ptr(dxC), _ptr(dxW), _ptr(dyC), _ptr(xc), _ptr(yc))
    return dict(dxC=dxC, dxW=dxW, dyC=dyC, xc=xc, yc=yc)


class Oracle:
    """FP64 CPU restatement of one solver configuration."""

    def __init__(self, nx, ny, domain=(0.0, 1.0, 0.0, 1.0), scheme=0, recon=0, flux=0, bp=0, bc=0,
                 rk=2, cfl=0.4, rho10=1.0, rho20=1.0, sound1=1.0, sound2=1.0, preb=0.0, gravity=0.0,
                 nranks=4, ghost=None):
        """bc: 0 periodic, 1 slip walls, 2 no-slip walls, 3 fixed ghost cells taken from
        `ghost`, a padded state (4, ny + 2*GHOST, nx + 2*GHOST)."""
        xL, xR, yL, yR = domain
        self.geo = geometry(nx, ny, xL, xR, yL, yR, nranks)
        self.ghost = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.float64)
        if bc == 3:
            assert self.ghost is not None and self.ghost.shape == (4, ny + 2 * GHOST, nx + 2 * GHOST)
        self.p = Params(nx=nx, ny=ny, scheme=scheme, recon=recon, flux=flux, bp=bp, bc=bc, rk=rk, cfl=cfl,
                        dX=(xR - xL) / nx, dY=(yR - yL) / ny, rho10=rho10, rho20=rho20, sound1=sound1,
                        sound2=sound2, preb=preb, gravity=gravity,
                        dxC=_ptr(self.geo["dxC"]), dxW=_ptr(self.geo["dxW"]), dyC=_ptr(self.geo["dyC"]),
                        ghost=_ptr(self.ghost) if self.ghost is not None else None)
        self.nx, self.ny = nx, ny

    # point kernels -------------------------------------------------------------------
    def volume_frac(self, U):
        return lib().orc_volume_frac(ctypes.byref(self.p), _ptr(_vec(U, 4)))

    def pressure(self, U):
        return lib().orc_pressure(ctypes.byref(self.p), _ptr(_vec(U, 4)))

    def sound_speed(self, U):
        return lib().orc_sound_speed(ctypes.byref(self.p), _ptr(_vec(U, 4)))

    def max_eig(self, U, nx, ny):
        return lib().orc_max_eig(ctypes.byref(self.p), _ptr(_vec(U, 4)), nx, ny)

    def phy_flux(self, U, nx, ny):
        F = np.zeros(4)
        lib().orc_phy_flux(ctypes.byref(self.p), _ptr(_vec(U, 4)), _ptr(F), nx, ny)
        return F

    def riemann(self, UL, UR, nx, ny, ftype):
        F = np.zeros(4)
        lib().orc_riemann(ctypes.byref(self.p), _ptr(_vec(UL, 4)), _ptr(_vec(UR, 4)), _ptr(F), nx, ny, ftype)
        return F

    def eigen(self, U, nx, ny):
        L, R = np.zeros(16), np.zeros(16)
        u = _vec(U, 4)
        lib().orc_eigen_L(ctypes.byref(self.p), _ptr(u), nx, ny, _ptr(L))
        lib().orc_eigen_R(ctypes.byref(self.p), _ptr(u), nx, ny, _ptr(R))
        return L.reshape(4, 4), R.reshape(4, 4)

    def bp_theta(self, U, faces):
        return lib().orc_bp_theta(ctypes.byref(self.p), _ptr(_vec(U, 4)), _ptr(_vec(faces, 64)))

    # grid operations -----------------------------------------------------------------
    def _state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4, self.ny, self.nx)
        return U

    def assemble_rhs(self, U, want_faces=False):
        U = self._state(U)
        rhs = np.zeros_like(U)
        faces = np.zeros((self.ny, self.nx, 64)) if want_faces else None
        theta = np.zeros((self.ny, self.nx)) if want_faces else None
        err = lib().orc_assemble_rhs(ctypes.byref(self.p), _ptr(U), _ptr(rhs), _ptr(faces), _ptr(theta))
        return rhs, faces, theta, err

    def compute_dt(self, U):
        dt = ctypes.c_double()
        err = lib().orc_compute_dt(ctypes.byref(self.p), _ptr(self._state(U)), ctypes.byref(dt))
        return dt.value, err

    def rk_step(self, U, dt):
        U = self._state(U).copy()
        err = lib().orc_rk_step(ctypes.byref(self.p), _ptr(U), dt)
        return U, err

    def run(self, U, t_end=0.0, fixed_steps=0, t_start=0.0, max_dts=100000):
        U = self._state(U).copy()
        dts = np.zeros(max_dts)
        n = ctypes.c_int()
        t = ctypes.c_double()
        err = lib().orc_run(ctypes.byref(self.p), _ptr(U), t_start, t_end, fixed_steps, _ptr(dts), max_dts,
                            ctypes.byref(n), ctypes.byref(t))
        return U, dts[: min(n.value, max_dts)], t.value, err


def weno(scheme, s, gp):
    return lib().orc_weno(scheme, _ptr(_vec(s, 5)), gp)


def weno7(scheme, s, gp, dx):
    """useWENO on a 7-point stencil (SCHEME 0, 1, 2)."""
    return lib().orc_weno7(scheme, _ptr(_vec(s, 7)), gp, dx)


def weno_exp(s, gp, dx):
    return lib().orc_weno_exp(_ptr(_vec(s, 7)), gp, dx)


def weno_exp_thinc(s, gp, dx):
    return lib().orc_weno_exp_thinc(_ptr(_vec(s, 7)), gp, dx)
