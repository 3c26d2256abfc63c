"""Python binding of libwenofv.so (the C ABI declared in include/wenofv.h).

The binding mirrors the reference's CWENOFV surface (include/WENOFV.hpp:108-257): a solver
is configured by the config/settings.cfg keys (TESTCASE, ELEMNUMX, ELEMNUMY, RKMETH, CFL,
SCHEME, PPlimiter, RIEMANNFLUX, RECONSTRUCTIONTYPE) and exposes calculateDeltaT,
performTimeIntegration, assembleRHS and the run loop.  State arrays are float64 numpy
arrays of shape (4, ny, nx) holding (z1*rho1, z2*rho2, rho*u, rho*v).

Loading fails loudly when the library has not been built: there is no fallback.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WENOFV_LIB", os.path.join(HERE, "libwenofv.so"))

OK, ERR_CONFIG, ERR_HIP, ERR_NONFINITE_DT, ERR_ARG = 0, 1, 2, 3, 4
GHOST = 4  # WFV_GHOST: ghost layers of a fixed-ghost boundary (bc 3)
# boundary kinds: 0 periodic, 1 slip walls (NEUMANN, FREESLIP), 2 no-slip walls (NoSlip),
# 3 fixed ghost cells (NEUMANN2; set them with set_ghost_cells)
FLAG_NAN_FLUX, FLAG_BAD_THETA, FLAG_NAN_DT = 1, 2, 4

_dp = ctypes.POINTER(ctypes.c_double)


class Options(ctypes.Structure):
    _fields_ = [("elem_num_x", ctypes.c_int), ("elem_num_y", ctypes.c_int), ("rk_method", ctypes.c_int),
                ("cfl", ctypes.c_double), ("scheme", ctypes.c_int), ("pp_limiter", ctypes.c_int),
                ("riemann_flux", ctypes.c_int), ("reconstruction", ctypes.c_int), ("precision", ctypes.c_int),
                ("boundary", ctypes.c_int), ("geometry_ranks", ctypes.c_int), ("rows_per_strip", ctypes.c_int)]


class Case(ctypes.Structure):
    _fields_ = [("xL", ctypes.c_double), ("xR", ctypes.c_double), ("yL", ctypes.c_double), ("yR", ctypes.c_double),
                ("rho10", ctypes.c_double), ("rho20", ctypes.c_double), ("sound1", ctypes.c_double),
                ("sound2", ctypes.c_double), ("preb", ctypes.c_double), ("gravity", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("rows_per_strip", ctypes.c_int), ("persistent", ctypes.c_int), ("stage_ms", ctypes.c_double)]


class WenofvError(RuntimeError):
    pass


def _cache_path():
    return os.environ.get("WENOFV_TUNING_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "wenofv_tuning.json"))


def _cache_load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _cache_store(key, value):
    d = _cache_load()
    d[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


_lib = None


def lib():
    """Load libwenofv.so; raises if it is missing (build it with build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WenofvError(f"{LIB_PATH} not built: run python build.py in {HERE}")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.wfv_create.argtypes = [ctypes.POINTER(Options), ctypes.POINTER(Case), ctypes.POINTER(vp)]
    L.wfv_destroy.argtypes = [vp]
    L.wfv_last_error.restype = ctypes.c_char_p
    L.wfv_last_error.argtypes = [vp]
    for n in ("wfv_set_state", "wfv_get_state", "wfv_set_state_device", "wfv_get_state_device"):
        getattr(L, n).argtypes = [vp, vp]
    L.wfv_compute_dt.argtypes = [vp, _dp]
    L.wfv_step.argtypes = [vp, ctypes.c_double]
    L.wfv_assemble_rhs.argtypes = [vp, _dp, _dp, _dp]
    L.wfv_advance.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              _dp, _dp, ctypes.c_int]
    L.wfv_get_flags.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
    L.wfv_synchronize.argtypes = [vp]
    L.wfv_set_stream.argtypes = [vp, vp]
    L.wfv_grid.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp]
    L.wfv_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Tuning)]
    L.wfv_set_ghost_cells.argtypes = [vp, _dp]
    L.wfv_set_rows_per_strip.argtypes = [vp, ctypes.c_int]
    L.wfv_get_rows_per_strip.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.wfv_record_limiter_flags.argtypes = [vp, ctypes.c_int]
    L.wfv_get_limiter_flags.argtypes = [vp, vp]
    L.wfv_advance_fixed_noalloc.argtypes = [ctypes.POINTER(Options), ctypes.POINTER(Case), vp, ctypes.c_double,
                                            ctypes.c_int, vp, ctypes.POINTER(ctypes.c_int), _dp]
    L.wfv_version.restype = ctypes.c_char_p
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Solver:
    """One single-domain WENOFV solver instance on the current GPU."""

    def __init__(self, nx, ny, domain=(0.0, 1.0, 0.0, 1.0), rho10=1.0, rho20=1.0, sound1=1.0, sound2=1.0, preb=0.0,
                 gravity=0.0, scheme=0, recon=0, flux=0, bp=0, rk=2, cfl=0.4, precision=0, bc=0, geometry_ranks=4,
                 rows_per_strip=0):
        L = lib()
        self.nx, self.ny = nx, ny
        self.opt = Options(nx, ny, rk, cfl, scheme, bp, flux, recon, precision, bc, geometry_ranks, rows_per_strip)
        xL, xR, yL, yR = domain
        self.case = Case(xL, xR, yL, yR, rho10, rho20, sound1, sound2, preb, gravity)
        h = ctypes.c_void_p()
        rc = L.wfv_create(ctypes.byref(self.opt), ctypes.byref(self.case), ctypes.byref(h))
        if rc != OK:
            raise WenofvError(f"wfv_create failed ({rc}): {L.wfv_last_error(None).decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().wfv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise WenofvError(f"wenofv error {rc}: {lib().wfv_last_error(self._h).decode()}")

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4, self.ny, self.nx)
        self._check(lib().wfv_set_state(self._h, U.ctypes.data))

    def get_state(self):
        U = np.empty((4, self.ny, self.nx))
        self._check(lib().wfv_get_state(self._h, U.ctypes.data))
        return U

    def set_ghost_cells(self, Upad):
        """Ghost cells of a boundary-3 solver: a padded state (4, ny + 2*GHOST, nx + 2*GHOST)."""
        Upad = np.ascontiguousarray(Upad, dtype=np.float64)
        assert Upad.shape == (4, self.ny + 2 * GHOST, self.nx + 2 * GHOST)
        self._check(lib().wfv_set_ghost_cells(self._h, _p(Upad)))

    def set_state_device(self, ptr):
        self._check(lib().wfv_set_state_device(self._h, ctypes.c_void_p(ptr)))

    def get_state_device(self, ptr):
        self._check(lib().wfv_get_state_device(self._h, ctypes.c_void_p(ptr)))

    def compute_dt(self):
        dt = ctypes.c_double()
        self._check(lib().wfv_compute_dt(self._h, ctypes.byref(dt)))
        return dt.value

    def step(self, dt):
        self._check(lib().wfv_step(self._h, dt))

    def assemble_rhs(self, want_faces=False):
        rhs = np.empty((4, self.ny, self.nx))
        faces = np.empty((self.ny, self.nx, 64)) if want_faces else None
        theta = np.empty((self.ny, self.nx)) if want_faces else None
        self._check(lib().wfv_assemble_rhs(self._h, _p(rhs), _p(faces), _p(theta)))
        return rhs, faces, theta

    def advance(self, t_end=0.0, fixed_steps=0, t_start=0.0, max_dts=0):
        n = ctypes.c_int()
        t = ctypes.c_double()
        dts = np.zeros(max_dts) if max_dts else None
        rc = lib().wfv_advance(self._h, t_start, t_end, fixed_steps, ctypes.byref(n), ctypes.byref(t), _p(dts),
                               max_dts)
        self._check(rc)
        return n.value, t.value, (dts[: min(n.value, max_dts)] if max_dts else None)

    def flags(self):
        f = ctypes.c_uint()
        self._check(lib().wfv_get_flags(self._h, ctypes.byref(f)))
        return f.value

    def set_stream(self, stream_handle):
        """Run on a caller-owned HIP stream (e.g. torch.cuda.Stream().cuda_stream); 0 resets."""
        self._check(lib().wfv_set_stream(self._h, ctypes.c_void_p(stream_handle or None)))

    def synchronize(self):
        self._check(lib().wfv_synchronize(self._h))

    def _key(self):
        o = self.opt
        return (f"solver:{version()}:p{o.precision}r{o.reconstruction}s{o.scheme}f{o.riemann_flux}b{o.pp_limiter}"
                f":{self.nx}x{self.ny}")

    def autotune(self, trials=2, cache=True):
        """Choose the rows per wavefront strip by timing stage launches on the current state
        (left unchanged).  With cache, an earlier choice for the same options and grid is
        reused.  Returns {"rows_per_strip", "stage_ms"}."""
        key = self._key()
        if cache:
            hit = _cache_load().get(key)
            if hit:
                self.set_rows_per_strip(hit["rows_per_strip"])
                return hit
        t = Tuning()
        self._check(lib().wfv_autotune(self._h, trials, ctypes.byref(t)))
        res = {"rows_per_strip": t.rows_per_strip, "stage_ms": t.stage_ms}
        if cache:
            _cache_store(key, res)
        return res

    def set_rows_per_strip(self, rows):
        self._check(lib().wfv_set_rows_per_strip(self._h, rows))

    def rows_per_strip(self):
        r = ctypes.c_int()
        self._check(lib().wfv_get_rows_per_strip(self._h, ctypes.byref(r)))
        return r.value

    def grid(self):
        nx, ny = ctypes.c_int(), ctypes.c_int()
        dxC, dxW, dyC = np.empty(self.nx), np.empty(self.nx), np.empty(self.ny)
        self._check(lib().wfv_grid(self._h, ctypes.byref(nx), ctypes.byref(ny), _p(dxC), _p(dxW), _p(dyC)))
        return dict(dxC=dxC, dxW=dxW, dyC=dyC)


def version():
    return lib().wfv_version().decode()


class Instance(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("boundary", ctypes.c_int), ("pad", ctypes.c_int),
                ("cs", Case)]


def _batch_lib():
    L = lib()
    if getattr(L, "_batch_ready", False):
        return L
    vp = ctypes.c_void_p
    L.wfv_batch_create.argtypes = [ctypes.POINTER(Options), ctypes.POINTER(Instance), ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.wfv_batch_destroy.argtypes = [vp]
    L.wfv_batch_last_error.restype = ctypes.c_char_p
    L.wfv_batch_last_error.argtypes = [vp]
    L.wfv_batch_set_state.argtypes = [vp, ctypes.c_int, vp]
    L.wfv_batch_get_state.argtypes = [vp, ctypes.c_int, vp]
    L.wfv_batch_advance.argtypes = [vp, ctypes.c_double, ctypes.c_int, vp, vp, vp]
    L.wfv_batch_set_stream.argtypes = [vp, vp]
    L.wfv_batch_set_rows_per_strip.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.wfv_batch_cells.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.wfv_batch_get_rows_per_strip.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.wfv_batch_set_schedule.argtypes = [vp, ctypes.c_int]
    L.wfv_batch_get_schedule.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.wfv_batch_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Tuning)]
    L.wfv_batch_set_ghost_cells.argtypes = [vp, ctypes.c_int, _dp]
    L._batch_ready = True
    return L


class Batch:
    """Independent instances with uneven grids advanced together (wfv_batch_* ABI).

    instances: list of dicts with nx, ny, bc and the case (domain, rho10, rho20, sound1,
    sound2, preb, gravity); options shared by the batch as keyword arguments."""

    def __init__(self, instances, scheme=0, recon=0, flux=0, bp=0, rk=2, cfl=0.4, precision=0, geometry_ranks=1,
                 rows_per_strip=0):
        L = _batch_lib()
        self.n = len(instances)
        self.shapes = [(d["ny"], d["nx"]) for d in instances]
        self.opt = Options(0, 0, rk, cfl, scheme, bp, flux, recon, precision, 0, geometry_ranks, rows_per_strip)
        arr = (Instance * self.n)()
        for k, d in enumerate(instances):
            xL, xR, yL, yR = d.get("domain", (0.0, 1.0, 0.0, 1.0))
            arr[k] = Instance(d["nx"], d["ny"], d.get("bc", 0), 0,
                              Case(xL, xR, yL, yR, d["rho10"], d["rho20"], d["sound1"], d["sound2"],
                                   d.get("preb", 0.0), d.get("gravity", 0.0)))
        h = ctypes.c_void_p()
        rc = L.wfv_batch_create(ctypes.byref(self.opt), arr, self.n, ctypes.byref(h))
        if rc != OK:
            raise WenofvError(f"wfv_batch_create failed ({rc}): {L.wfv_batch_last_error(None).decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _batch_lib().wfv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise WenofvError(f"wenofv batch error {rc}: {_batch_lib().wfv_batch_last_error(self._h).decode()}")

    def set_state(self, i, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4,) + self.shapes[i]
        self._check(_batch_lib().wfv_batch_set_state(self._h, i, U.ctypes.data))

    def set_ghost_cells(self, i, Upad):
        ny, nx = self.shapes[i]
        Upad = np.ascontiguousarray(Upad, dtype=np.float64)
        assert Upad.shape == (4, ny + 2 * GHOST, nx + 2 * GHOST)
        self._check(_batch_lib().wfv_batch_set_ghost_cells(self._h, i, _p(Upad)))

    def get_state(self, i):
        U = np.empty((4,) + self.shapes[i])
        self._check(_batch_lib().wfv_batch_get_state(self._h, i, U.ctypes.data))
        return U

    def advance(self, t_end=0.0, fixed_steps=0):
        n = np.zeros(self.n, dtype=np.int32)
        t = np.zeros(self.n)
        f = np.zeros(self.n, dtype=np.uint32)
        self._check(_batch_lib().wfv_batch_advance(self._h, t_end, fixed_steps, n.ctypes.data, t.ctypes.data,
                                                   f.ctypes.data))
        return n, t, f

    def set_stream(self, stream_handle):
        self._check(_batch_lib().wfv_batch_set_stream(self._h, ctypes.c_void_p(stream_handle or None)))

    def set_rows_per_strip(self, i, rows):
        self._check(_batch_lib().wfv_batch_set_rows_per_strip(self._h, i, rows))

    def rows_per_strip(self, i):
        r = ctypes.c_int()
        self._check(_batch_lib().wfv_batch_get_rows_per_strip(self._h, i, ctypes.byref(r)))
        return r.value

    def set_schedule(self, persistent):
        self._check(_batch_lib().wfv_batch_set_schedule(self._h, int(persistent)))

    def schedule(self):
        p = ctypes.c_int()
        self._check(_batch_lib().wfv_batch_get_schedule(self._h, ctypes.byref(p)))
        return p.value

    def _key(self):
        o = self.opt
        classes = {}
        for ny, nx in self.shapes:
            classes[f"{nx}x{ny}"] = classes.get(f"{nx}x{ny}", 0) + 1
        sig = ",".join(f"{k}*{v}" for k, v in sorted(classes.items()))
        return f"batch:{version()}:p{o.precision}r{o.reconstruction}s{o.scheme}f{o.riemann_flux}b{o.pp_limiter}:{sig}"

    def autotune(self, trials=2, cache=True):
        """Choose rows per strip per grid-size class and the launch schedule by timing the
        one-launch stage on the current states (left unchanged).  With cache, an earlier
        choice for the same options and class mix is reused.  Returns {"rows_per_strip":
        {"NXxNY": rows}, "persistent": 0|1, "stage_ms"}."""
        key = self._key()
        hit = _cache_load().get(key) if cache else None
        if hit:
            for i, (ny, nx) in enumerate(self.shapes):
                self.set_rows_per_strip(i, hit["rows_per_strip"][f"{nx}x{ny}"])
            self.set_schedule(hit["persistent"])
            return hit
        t = Tuning()
        self._check(_batch_lib().wfv_batch_autotune(self._h, trials, ctypes.byref(t)))
        rows = {}
        for i, (ny, nx) in enumerate(self.shapes):
            rows[f"{nx}x{ny}"] = self.rows_per_strip(i)
        res = {"rows_per_strip": rows, "persistent": t.persistent, "stage_ms": t.stage_ms}
        if cache:
            _cache_store(key, res)
        return res

    def cells(self):
        c = ctypes.c_longlong()
        self._check(_batch_lib().wfv_batch_cells(self._h, ctypes.byref(c)))
        return c.value


def reference_case(testcase):
    """Parameters of a reference test case (TESTCASE 0..8): dict for Solver(**...), the
    boundary and the final output time (host only)."""
    L = lib()
    L.wfv_reference_case.argtypes = [ctypes.c_int, ctypes.POINTER(Case), ctypes.POINTER(ctypes.c_int), _dp]
    c, bc, t = Case(), ctypes.c_int(), ctypes.c_double()
    if L.wfv_reference_case(testcase, ctypes.byref(c), ctypes.byref(bc), ctypes.byref(t)) != OK:
        raise WenofvError(f"unknown test case {testcase}")
    kw = dict(domain=(c.xL, c.xR, c.yL, c.yR), rho10=c.rho10, rho20=c.rho20, sound1=c.sound1, sound2=c.sound2,
              preb=c.preb, gravity=c.gravity, bc=bc.value)
    return kw, t.value


def reference_initial_state_padded(testcase, nx, ny, geometry_ranks=4):
    """The same on the grid padded by GHOST cells per side (ghost cells of bc 3)."""
    L = lib()
    L.wfv_reference_initial_state_padded.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    U = np.empty((4, ny + 2 * GHOST, nx + 2 * GHOST))
    if L.wfv_reference_initial_state_padded(testcase, nx, ny, geometry_ranks, _p(U)) != OK:
        raise WenofvError("bad arguments")
    return U


def reference_initial_state(testcase, nx, ny, geometry_ranks=4):
    """Initial cell averages of a reference test case as CWENOFV::initializeAve builds them."""
    L = lib()
    L.wfv_reference_initial_state.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    U = np.empty((4, ny, nx))
    if L.wfv_reference_initial_state(testcase, nx, ny, geometry_ranks, _p(U)) != OK:
        raise WenofvError("bad arguments")
    return U
