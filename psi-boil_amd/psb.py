"""Python front end of libpsb.so, the MI355X pressure-Poisson solver.

Mirrors PSI-BOIL's solve surface (Src/Solver/Linear/linear.h:46-56):

    s = Solver(nx, ny, nz)             # CG(const Domain&, Prec::di())
    s.set_matrix(A)                    # the Matrix (c, w, e, s, n, b, t)
    x, res = s.solve(x, b, max_iter=..., res_rat=..., res_tol=...)

The library must be built (psi-boil_amd/build.py); importing this module on a
machine without it raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsb.so")

PSB_CG, PSB_BICGS = 0, 1
PSB_PREC_DIAG, PSB_PREC_AC, PSB_PREC_SGS = 0, 1, 2
PSB_FP64, PSB_MIXED = 0, 1
PSB_NORM_REF, PSB_NORM_L2REL = 0, 1
PSB_OK, PSB_MAXITER, PSB_STALE, PSB_BREAKDOWN, PSB_DIVERGED, PSB_NAN = 0, 1, 2, 3, 4, 5

STATUS = {0: "OK", 1: "MAXITER", 2: "STALE", 3: "BREAKDOWN", 4: "DIVERGED", 5: "NAN",
          -1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ENOTREADY"}


class PsbError(RuntimeError):
    pass


class Layout(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ghost", C.c_int)]


class Matrix7(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in "cwesnbt"]


class Stop(C.Structure):
    _fields_ = [("struct_size", C.c_int), ("min_iter", C.c_int), ("max_iter", C.c_int),
                ("res_rat", C.c_double), ("res_tol", C.c_double), ("scale", C.c_double),
                ("stalecount", C.c_int), ("precform", C.c_int), ("method", C.c_int), ("precond", C.c_int),
                ("precision", C.c_int), ("norm", C.c_int), ("ac_pre", C.c_int),
                ("ac_post", C.c_int), ("ac_omega", C.c_double), ("ac_cfac", C.c_double),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int), ("graph_chunk", C.c_int),
                ("inner_adapt", C.c_int), ("ac_coarsen", C.c_int),
                ("ac_store", C.c_int), ("ac_fuse", C.c_int), ("ac_kcycle", C.c_int),
                ("ac_flexible", C.c_int)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("converged", C.c_int),
                ("status", C.c_int), ("res0", C.c_double), ("res", C.c_double),
                ("rel_l2", C.c_double)]

    def asdict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["status_name"] = STATUS.get(self.status, "?")
        return d


KERNEL_KINDS = ["sweep", "sweep_prol", "sweep_last", "restrict", "cg_pq", "cg_xr", "ru_resid",
                "ru_xupdate"]


class KernelStat(C.Structure):
    _fields_ = [("launches", C.c_int), ("ms", C.c_double), ("bytes", C.c_double)]


class Metrics(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("dxc", "dxw", "dxe", "dyc", "dys", "dyn", "dzc",
                                          "dzb", "dzt")]


PSB_CYCLE_NONE, PSB_CYCLE_V, PSB_CYCLE_F, PSB_CYCLE_W, PSB_CYCLE_FLEX = -1, 0, 1, 2, 3


class Cycle(C.Structure):
    _fields_ = [("struct_size", C.c_int), ("res_tol", C.c_double), ("res_rat", C.c_double),
                ("max_cycles", C.c_int),
                ("min_cycles", C.c_int), ("max_iter", C.c_int * 3), ("dt", C.c_double),
                ("loop", C.c_int), ("init", C.c_int), ("stop_if_div", C.c_int),
                ("priority_min_cyc", C.c_int), ("use_linf", C.c_int),
                ("smooth_res_rat", C.c_double), ("smooth_res_tol", C.c_double),
                ("max_stale", C.c_int * 3), ("max_visits", C.c_int), ("level_solver", C.c_int)]


def _cycle(**cyc):
    c = Cycle()
    lib().psb_cycle_defaults(C.byref(c))
    shapes = {None: PSB_CYCLE_NONE, "V": PSB_CYCLE_V, "F": PSB_CYCLE_F, "W": PSB_CYCLE_W,
              "flex": PSB_CYCLE_FLEX}
    for k, v in cyc.items():
        if k in ("max_iter", "max_stale"):
            getattr(c, k)[:] = list(v)
        elif k in ("loop", "init") and (v is None or isinstance(v, str)):
            setattr(c, k, shapes[v])
        elif k == "level_solver" and isinstance(v, str):
            c.level_solver = {"cg": PSB_CG, "bicgs": PSB_BICGS}[v]
        else:
            setattr(c, k, v)
    return c


class Case(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nsph", C.c_int),
                ("rmin", C.c_double), ("rmax", C.c_double), ("rho_in", C.c_double),
                ("rho_out", C.c_double), ("seed_sph", C.c_ulonglong),
                ("seed_vel", C.c_ulonglong), ("centred", C.c_int), ("cx", C.c_double),
                ("cy", C.c_double), ("cz", C.c_double), ("stretch_z", C.c_double),
                ("lz", C.c_double), ("cylinder", C.c_int), ("cyl_d", C.c_double)]


class Bnd(C.Structure):
    _fields_ = [("type", C.c_int * 6), ("value", C.c_double * 6)]


EXPORTS = ["psb_stop_defaults", "psb_status_string", "psb_create", "psb_destroy",
           "psb_set_geometry", "psb_set_matrix", "psb_set_active", "psb_solve",
           "psb_solve_device", "psb_apply", "psb_history", "psb_generate",
           "psb_time_kernel", "psb_batch_cases", "psb_batch_create", "psb_batch_destroy",
           "psb_batch_system", "psb_batch_set_tuning", "psb_batch_plan", "psb_batch_solve",
           "psb_batch_solve_device", "psb_cycle_defaults", "psb_ac_vcycle",
           "psb_ac_vcycle_device", "psb_assemble_pressure", "psb_assemble_pressure_device",
           "psb_update_rhs_device", "psb_project_device", "psb_bck_name", "psb_save_bck",
           "psb_load_bck", "psb_save_bck_device", "psb_load_bck_device",
           "psb_profile_solve_device", "psb_solve_device_static", "psb_pressure_ghost_device",
           "psb_set_boundary"]

_lib = None


def lib():
    """Load libpsb.so; raises PsbError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PsbError(f"{LIB_PATH} is missing: build it with psi-boil_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.psb_status_string.restype = C.c_char_p
        L.psb_batch_system.restype = C.c_void_p
        for name in EXPORTS:
            getattr(L, name)
        _lib = L
    return _lib


def _dp(a):
    return C.c_void_p(a.ctypes.data)


def stop_defaults(**kw):
    s = Stop()
    lib().psb_stop_defaults(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def _check(rc, what):
    if rc < 0:
        raise PsbError(f"{what}: {STATUS.get(rc, rc)} ({lib().psb_status_string(rc).decode()})")
    return rc


class Solver:
    """One solver context (device buffers, hierarchy, stream, graphs)."""

    def __init__(self, nx, ny, nz, ghost=0, _ctx=None):
        self.shape = (nx, ny, nz)
        self.ghost = ghost
        self._owned = _ctx is None
        if _ctx is not None:  # a context owned by a Batch
            self._c = C.c_void_p(_ctx)
            return
        self._c = C.c_void_p()
        _check(lib().psb_create(C.byref(self._c), C.byref(Layout(nx, ny, nz, ghost))),
               "psb_create")

    def close(self):
        if self._c and self._owned:
            lib().psb_destroy(self._c)
        self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def host_size(self):
        g = 2 * self.ghost
        return (self.shape[0] + g) * (self.shape[1] + g) * (self.shape[2] + g)

    def _host(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.host_size:
            raise ValueError(f"array of {a.size} values, layout needs {self.host_size}")
        return a

    def set_geometry(self, dxc=None, dyc=None, dzc=None):
        if dxc is None:
            return _check(lib().psb_set_geometry(self._c, None, None, None), "psb_set_geometry")
        self._geo = [np.ascontiguousarray(v, dtype=np.float64) for v in (dxc, dyc, dzc)]
        return _check(lib().psb_set_geometry(self._c, *[_dp(v) for v in self._geo]),
                      "psb_set_geometry")

    def set_active(self, off):
        self._off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
        return _check(lib().psb_set_active(self._c, _dp(self._off)), "psb_set_active")

    def set_matrix(self, A):
        arrs = [self._host(A[k]) for k in "cwesnbt"]
        m = Matrix7(*[a.ctypes.data for a in arrs])
        return _check(lib().psb_set_matrix(self._c, C.byref(m)), "psb_set_matrix")

    def solve(self, x, b, **stop):
        x = np.array(self._host(x), copy=True)
        b = self._host(b)
        s = stop_defaults(**stop)
        r = Result()
        _check(lib().psb_solve(self._c, _dp(x), _dp(b), C.byref(s), C.byref(r)), "psb_solve")
        return x, r.asdict()

    def solve_device(self, d_x, d_b, **stop):
        """d_x, d_b: device pointers (int) to compact FP64 arrays."""
        s = stop_defaults(**stop)
        r = Result()
        _check(lib().psb_solve_device(self._c, C.c_void_p(d_x), C.c_void_p(d_b), C.byref(s),
                                      C.byref(r)), "psb_solve_device")
        return r.asdict()

    def solve_device_static(self, d_x, d_b, **stop):
        """solve_device for CG-di in FP64 through the entry point that
        allocates nothing (its work arrays in the library's static device
        array; at most 2^21 cells)."""
        s = stop_defaults(**stop)
        r = Result()
        _check(lib().psb_solve_device_static(self._c, C.c_void_p(d_x), C.c_void_p(d_b),
                                             C.byref(s), C.byref(r)),
               "psb_solve_device_static")
        return r.asdict()

    def profile_solve_device(self, d_x, d_b, **stop):
        """One solve as solve_device runs it, launched kernel by kernel with
        the finest level's kernels timed by HIP events: (result, {kind:
        {launches, ms, bytes}}, solve_ms)."""
        s = stop_defaults(**stop)
        r = Result()
        st = (KernelStat * len(KERNEL_KINDS))()
        ms = C.c_double()
        _check(lib().psb_profile_solve_device(self._c, C.c_void_p(d_x), C.c_void_p(d_b),
                                              C.byref(s), C.byref(r), st, C.byref(ms)),
               "psb_profile_solve_device")
        stats = {k: {"launches": st[q].launches, "ms": st[q].ms, "bytes": st[q].bytes}
                 for q, k in enumerate(KERNEL_KINDS)}
        return r.asdict(), stats, ms.value

    def assemble_pressure(self, rho, mx, my, mz, off=None):
        """Pressure::discretize on the device.  rho with one ghost layer;
        mx = (dxc, dxw, dxe), my = (dyc, dys, dyn), mz = (dzc, dzb, dzt)."""
        self._met = [np.ascontiguousarray(a, dtype=np.float64) for a in (*mx, *my, *mz)]
        m = Metrics(*[a.ctypes.data for a in self._met])
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        offp = None
        if off is not None:
            self._off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
            offp = _dp(self._off)
        return _check(lib().psb_assemble_pressure(self._c, _dp(rho), C.byref(m), offp),
                      "psb_assemble_pressure")

    def assemble_pressure_device(self, d_rho, mx, my, mz, off=None):
        """As assemble_pressure with the ghosted density already on the device."""
        self._met = [np.ascontiguousarray(a, dtype=np.float64) for a in (*mx, *my, *mz)]
        m = Metrics(*[a.ctypes.data for a in self._met])
        offp = None
        if off is not None:
            self._off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
            offp = _dp(self._off)
        return _check(lib().psb_assemble_pressure_device(self._c, C.c_void_p(d_rho), C.byref(m),
                                                         offp), "psb_assemble_pressure_device")

    def update_rhs_device(self, d_u, d_v, d_w, d_b, zero_mean=True):
        return _check(lib().psb_update_rhs_device(self._c, C.c_void_p(d_u), C.c_void_p(d_v),
                                                  C.c_void_p(d_w), int(zero_mean),
                                                  C.c_void_p(d_b)), "psb_update_rhs_device")

    def project_device(self, d_rho, d_p, d_u, d_v, d_w, dt=1.0, mask=7):
        """Momentum::project on device face velocities (in place)."""
        return _check(lib().psb_project_device(self._c, C.c_void_p(d_rho), C.c_void_p(d_p),
                                               C.c_double(dt), int(mask), C.c_void_p(d_u),
                                               C.c_void_p(d_v), C.c_void_p(d_w)),
                      "psb_project_device")

    def set_boundary(self, types, values):
        """The pressure's boundary condition per side (imin .. kmax): types 0
        Neumann group, 1 Dirichlet group; used by the next assembly."""
        b = Bnd((C.c_int * 6)(*[int(t) for t in types]),
                (C.c_double * 6)(*[float(v) for v in values]))
        return _check(lib().psb_set_boundary(self._c, C.byref(b)), "psb_set_boundary")

    def pressure_ghost_device(self, d_c, d_kappa, d_rho, sigma, d_fext):
        """Pressure::ghost: d_fext += the ghost-fluid surface-tension source
        (d_c, d_kappa, d_rho ghosted as the assembly's density)."""
        return _check(lib().psb_pressure_ghost_device(self._c, C.c_void_p(d_c),
                                                      C.c_void_p(d_kappa), C.c_void_p(d_rho),
                                                      C.c_double(sigma), C.c_void_p(d_fext)),
                      "psb_pressure_ghost_device")

    def save_bck_device(self, path, d_x):
        return _check(lib().psb_save_bck_device(self._c, path.encode(), C.c_void_p(d_x)),
                      "psb_save_bck_device")

    def load_bck_device(self, path, d_x):
        return _check(lib().psb_load_bck_device(self._c, path.encode(), C.c_void_p(d_x)),
                      "psb_load_bck_device")

    def ac_vcycle(self, x, b, **cyc):
        """The reference's AC::cycle (CG smoothers on its own levels); loop and
        init may be given as "V", "F", "W" (init None: no full-cycle start)."""
        x = np.array(self._host(x), copy=True)
        b = self._host(b)
        c = _cycle(**cyc)
        r = Result()
        _check(lib().psb_ac_vcycle(self._c, _dp(x), _dp(b), C.byref(c), C.byref(r)),
               "psb_ac_vcycle")
        return x, r.asdict()

    def ac_vcycle_device(self, d_x, d_b, **cyc):
        c = _cycle(**cyc)
        r = Result()
        _check(lib().psb_ac_vcycle_device(self._c, C.c_void_p(d_x), C.c_void_p(d_b),
                                          C.byref(c), C.byref(r)), "psb_ac_vcycle_device")
        return r.asdict()

    def apply(self, x, b=None):
        x = self._host(x)
        y = np.zeros(self.host_size)
        bp = _dp(self._host(b)) if b is not None else None
        _check(lib().psb_apply(self._c, _dp(x), bp, _dp(y)), "psb_apply")
        return y

    def history(self, cap=100000):
        h = np.zeros(cap)
        n = _check(lib().psb_history(self._c, _dp(h), cap), "psb_history")
        return h[:n]

    def time_kernel(self, kernel, reps=20):
        ms, by = C.c_double(), C.c_double()
        _check(lib().psb_time_kernel(self._c, kernel, reps, C.byref(ms), C.byref(by)),
               "psb_time_kernel")
        return ms.value, by.value


def generate(nx, ny, nz, nsph=0, rmin=0.0, rmax=0.0, rho_in=1.0, rho_out=1.0, seed_sph=0,
             seed_vel=0, centre=None, stretch_z=0.0, lz=0.0, cylinder=False, cyl_d=0.2):
    """Synthetic pressure system; returns dict with c..t, rhs, off, dxc, dyc, dzc."""
    cs = Case(nx, ny, nz, nsph, rmin, rmax, rho_in, rho_out, seed_sph, seed_vel,
              1 if centre is not None else 0, *(centre if centre is not None else (0, 0, 0)),
              stretch_z, lz, 1 if cylinder else 0, cyl_d)
    N = nx * ny * nz
    out = {k: np.zeros(N) for k in "cwesnbt"}
    out["rhs"] = np.zeros(N)
    out["off"] = np.zeros(N, dtype=np.uint8)
    out["dxc"], out["dyc"], out["dzc"] = np.zeros(nx), np.zeros(ny), np.zeros(nz)
    _check(lib().psb_generate(C.byref(cs), *[_dp(out[k]) for k in "cwesnbt"], _dp(out["rhs"]),
                              _dp(out["off"]), _dp(out["dxc"]), _dp(out["dyc"]),
                              _dp(out["dzc"])), "psb_generate")
    return out


# --------------------------------------------------------------------------
# device buffers through the HIP runtime (plumbing for tests and benchmarks)
# --------------------------------------------------------------------------
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipGetErrorString.restype = C.c_char_p
    return _hip


def _hipchk(rc, what):
    if rc != 0:
        raise PsbError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A float64 device array (hipMalloc) with host copies in and out."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = C.c_void_p()
        _hipchk(hip().hipMalloc(C.byref(self.ptr), C.c_size_t(8 * max(self.n, 1))), "hipMalloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        _hipchk(hip().hipMemcpy(self.ptr, C.c_void_p(a.ctypes.data), C.c_size_t(8 * self.n), 1),
                "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n)
        _hipchk(hip().hipMemcpy(C.c_void_p(out.ctypes.data), self.ptr, C.c_size_t(8 * self.n), 2),
                "hipMemcpy D2H")
        return out

    def zero(self):
        _hipchk(hip().hipMemset(self.ptr, 0, C.c_size_t(8 * self.n)), "hipMemset")

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bck_name(name, step=-1, rank=-1):
    """name_file(name, ".bck", step, rank) of the reference."""
    buf = C.create_string_buffer(4096)
    _check(lib().psb_bck_name(name.encode(), int(step), int(rank), buf, 4096), "psb_bck_name")
    return buf.value.decode()


def save_bck(path, x, shape, bw=3):
    """Scalar::save of a compact interior field x of the given (nx, ny, nz)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    _check(lib().psb_save_bck(path.encode(), *shape, int(bw), _dp(x)), "psb_save_bck")


def load_bck(path, shape):
    """Scalar::load into a compact interior field; returns (x, ghost width)."""
    x = np.zeros(int(np.prod(shape)))
    bw = C.c_int(0)
    _check(lib().psb_load_bck(path.encode(), *shape, _dp(x), C.byref(bw)), "psb_load_bck")
    return x, bw.value


def device_copy(dst, src):
    """dst <- src, two DeviceArrays of the same length (device to device)."""
    if dst.n != src.n:
        raise PsbError("device_copy: lengths differ")
    _hipchk(hip().hipMemcpy(dst.ptr, src.ptr, C.c_size_t(8 * src.n), 3), "hipMemcpy D2D")


def device_synchronize():
    _hipchk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


# --------------------------------------------------------------------------
# batched solves
# --------------------------------------------------------------------------
def batch_cases(seed, nsys):
    """The seeded batch of cases (list of dicts of generate() arguments)."""
    arr = (Case * nsys)()
    _check(lib().psb_batch_cases(C.c_ulonglong(seed), nsys, arr), "psb_batch_cases")
    out = []
    for c in arr:
        out.append(dict(nx=c.nx, ny=c.ny, nz=c.nz, nsph=c.nsph, rmin=c.rmin, rmax=c.rmax,
                        rho_in=c.rho_in, rho_out=c.rho_out, seed_sph=c.seed_sph,
                        seed_vel=c.seed_vel))
    return out


class Batch:
    """Many systems solved together (psb_batch_*).  systems[s] is a Solver view
    of system s's context for set_geometry / set_active / set_matrix."""

    def __init__(self, shapes, ghost=0):
        self.shapes = [tuple(sh) for sh in shapes]
        lays = (Layout * len(shapes))(*[Layout(*sh, ghost) for sh in self.shapes])
        self._b = C.c_void_p()
        _check(lib().psb_batch_create(C.byref(self._b), len(shapes), lays), "psb_batch_create")
        self.systems = [Solver(*sh, ghost=ghost,
                               _ctx=lib().psb_batch_system(self._b, q))
                        for q, sh in enumerate(self.shapes)]

    def close(self):
        if self._b:
            for s in self.systems:
                s.close()
            lib().psb_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, vpt=0):
        return _check(lib().psb_batch_set_tuning(self._b, vpt), "psb_batch_set_tuning")

    def plan(self, cap=1024):
        t = (C.c_int * (3 * cap))()
        n = _check(lib().psb_batch_plan(self._b, t, cap), "psb_batch_plan")
        return [tuple(t[3 * q:3 * q + 3]) for q in range(min(n, cap))]

    def _results(self, R):
        return [r.asdict() for r in R]

    def solve(self, xs, bs, **stop):
        n = len(self.shapes)
        xs = [np.array(s._host(x), copy=True) for s, x in zip(self.systems, xs)]
        bs = [s._host(b) for s, b in zip(self.systems, bs)]
        px = (C.c_void_p * n)(*[x.ctypes.data for x in xs])
        pb = (C.c_void_p * n)(*[b.ctypes.data for b in bs])
        st = stop_defaults(**dict(BATCH_STOP, **stop))
        R = (Result * n)()
        rc = _check(lib().psb_batch_solve(self._b, px, pb, C.byref(st), R), "psb_batch_solve")
        return xs, self._results(R), rc

    def solve_device(self, d_xs, d_bs, **stop):
        """d_xs, d_bs: device pointers (int) to compact FP64 arrays per system."""
        n = len(self.shapes)
        px = (C.c_void_p * n)(*d_xs)
        pb = (C.c_void_p * n)(*d_bs)
        st = stop_defaults(**dict(BATCH_STOP, **stop))
        R = (Result * n)()
        rc = _check(lib().psb_batch_solve_device(self._b, px, pb, C.byref(st), R),
                    "psb_batch_solve_device")
        return self._results(R), rc


# the batched path's method: FP32 AC-PCG with FP64 reliable updates
BATCH_STOP = dict(precond=PSB_PREC_AC, method=PSB_CG, precision=PSB_MIXED, inner_adapt=2,
                  inner_rtol=1e-2, norm=PSB_NORM_L2REL, res_tol=0.0)
