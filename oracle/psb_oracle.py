"""ctypes front end of the C oracle (oracle/psb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg.  The product library never imports it.

Arrays are numpy float64 in the interior layout [i][j][k], k fastest.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsb_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class OraMat(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + [
        (k, C.POINTER(C.c_double)) for k in "cwesnbt"
    ]


class OraResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("reason", C.c_int),
                ("res0", C.c_double), ("res", C.c_double)]


MAXL = 24


class OraCycle(C.Structure):
    _fields_ = [("toler", C.c_double), ("factor", C.c_double), ("max_cyc", C.c_int),
                ("min_cyc", C.c_int), ("mv", C.c_int * 3), ("ms", C.c_int * 3),
                ("dt", C.c_double),
                ("init", C.c_int), ("loop", C.c_int), ("stop_if_div", C.c_int),
                ("priority_min_cyc", C.c_int), ("use_linf", C.c_int), ("resrat", C.c_double),
                ("restol", C.c_double), ("max_visits", C.c_int), ("solver", C.c_int)]


class OraAC(C.Structure):
    _fields_ = [("nlev", C.c_int), ("nx", C.c_int * MAXL), ("ny", C.c_int * MAXL),
                ("nz", C.c_int * MAXL), ("A", OraMat * MAXL),
                ("aflag", C.POINTER(C.c_int) * MAXL),
                ("dV", C.POINTER(C.c_double) * MAXL),
                ("phi", C.POINTER(C.c_double) * MAXL),
                ("f", C.POINTER(C.c_double) * MAXL),
                ("res", C.POINTER(C.c_double) * MAXL),
                ("G", C.POINTER(C.c_double))]


def build():
    """Compile the oracle with gcc (no FMA contraction)."""
    src = os.path.join(HERE, "psb_oracle.c")
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC",
                           "-shared", "-o", LIB_PATH, src, "-lm"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_u01.restype = C.c_double
        L.ora_splitmix64.restype = C.c_uint64
        L.ora_dot.restype = C.c_double
        L.ora_dot_voldiv_avg.restype = C.c_double
        L.ora_gen_spheres.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, _dp]
        L.ora_metrics.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ora_face_velocity.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int,
                                        _dp, _dp, _dp]
        _lib = L
    return _lib


class Mat:
    """Seven coefficient arrays plus the ctypes view the C side takes."""

    def __init__(self, nx, ny, nz, arrays=None):
        self.shape = (nx, ny, nz)
        n = nx * ny * nz
        self.a = {}
        for k in "cwesnbt":
            v = np.zeros(n) if arrays is None else np.ascontiguousarray(arrays[k], dtype=np.float64).ravel()
            self.a[k] = v
        self.s = OraMat(nx, ny, nz, *[self.a[k].ctypes.data_as(C.POINTER(C.c_double)) for k in "cwesnbt"])

    @property
    def n(self):
        return int(np.prod(self.shape))

    def __getitem__(self, k):
        return self.a[k]


def _p(x):
    return x.ctypes.data_as(C.c_void_p)


def uniform_nodes(n, L=1.0):
    return np.array([L * i / n for i in range(n + 1)], dtype=np.float64)


def metrics(xn):
    n = len(xn) - 1
    xc, dxc, dxw, dxe = (np.zeros(n) for _ in range(4))
    lib().ora_metrics(n, np.ascontiguousarray(xn, dtype=np.float64), xc, dxc, dxw, dxe)
    return xc, dxc, dxw, dxe


def spheres(seed, nsph, rmin, rmax, L=(1.0, 1.0, 1.0)):
    out = np.zeros(4 * max(nsph, 1))
    lib().ora_gen_spheres(seed, nsph, rmin, rmax, L[0], L[1], L[2], out)
    return out[:4 * nsph].reshape(nsph, 4) if nsph else np.zeros((0, 4))


def density(xc, yc, zc, sph, rho_in, rho_out):
    nx, ny, nz = len(xc), len(yc), len(zc)
    rho = np.zeros((nx + 2) * (ny + 2) * (nz + 2))
    s = np.ascontiguousarray(sph, dtype=np.float64).ravel()
    lib().ora_density(nx, ny, nz, _p(xc), _p(yc), _p(zc), len(sph), _p(s) if len(s) else None,
                      C.c_double(rho_in), C.c_double(rho_out), _p(rho))
    return rho


def discretize(mx, my, mz, rho, off=None, bnd=None):
    """mx = (dxc, dxw, dxe) etc.  Returns (Mat, ci, dV); with bnd = (types,
    values) per side (imin .. kmax; 0 Neumann group, 1 Dirichlet group) also
    the boundary source fbnd."""
    nx, ny, nz = len(mx[0]), len(my[0]), len(mz[0])
    A = Mat(nx, ny, nz)
    ci = np.zeros(A.n)
    dV = np.zeros(A.n)
    offp = None
    if off is not None:
        off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
        offp = _p(off)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (*mx, *my, *mz)]
    if bnd is None:
        lib().ora_discretize(nx, ny, nz, *[_p(a) for a in arrs], _p(rho), offp,
                             C.byref(A.s), _p(ci), _p(dV))
        return A, ci, dV
    bt = (C.c_int * 6)(*[int(t) for t in bnd[0]])
    bv = (C.c_double * 6)(*[float(v) for v in bnd[1]])
    fbnd = np.zeros(A.n)
    lib().ora_discretize_bnd(nx, ny, nz, *[_p(a) for a in arrs], _p(rho), offp, bt, bv,
                             C.byref(A.s), _p(ci), _p(dV), _p(fbnd))
    return A, ci, dV, fbnd


def face_velocity(seed, nx, ny, nz):
    u = np.zeros((nx + 1) * ny * nz)
    v = np.zeros(nx * (ny + 1) * nz)
    w = np.zeros(nx * ny * (nz + 1))
    lib().ora_face_velocity(seed, nx, ny, nz, u, v, w)
    return u, v, w


def rhs(dxc, dyc, dzc, u, v, w, off=None, zero_mean=True, fbnd=None):
    nx, ny, nz = len(dxc), len(dyc), len(dzc)
    b = np.zeros(nx * ny * nz)
    offp = None
    if off is not None:
        off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
        offp = _p(off)
    fb = None if fbnd is None else _p(np.ascontiguousarray(fbnd, dtype=np.float64))
    lib().ora_rhs_bnd(nx, ny, nz, _p(dxc), _p(dyc), _p(dzc), _p(u), _p(v), _p(w), offp, fb,
                      int(zero_mean), _p(b))
    return b


def project(mx, my, mz, rho, p, u, v, w, dt=1.0, off=None, mask=7):
    """Momentum::project on copies of the face velocities; mx = (dxc, dxw, dxe)."""
    nx, ny, nz = len(mx[0]), len(my[0]), len(mz[0])
    u, v, w = (np.array(a, dtype=np.float64) for a in (u, v, w))
    offp = None
    if off is not None:
        off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
        offp = _p(off)
    lib().ora_project(nx, ny, nz, _p(np.ascontiguousarray(mx[1], dtype=np.float64)),
                      _p(np.ascontiguousarray(my[1], dtype=np.float64)),
                      _p(np.ascontiguousarray(mz[1], dtype=np.float64)), _p(rho), offp,
                      _p(np.ascontiguousarray(p, dtype=np.float64)), C.c_double(dt), int(mask),
                      _p(u), _p(v), _p(w))
    return u, v, w


def ghost(mx, my, mz, c, kappa, rho, sigma, fext=None, off=None):
    """Pressure::ghost: fext (a copy; zeros if None) plus the ghost-fluid
    surface-tension source; mx = (dxc, dxw, dxe) etc., c, kappa, rho with
    their ghost layer."""
    nx, ny, nz = len(mx[0]), len(my[0]), len(mz[0])
    f = np.zeros(nx * ny * nz) if fext is None else np.array(fext, dtype=np.float64)
    offp = None
    if off is not None:
        off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
        offp = _p(off)
    a = [np.ascontiguousarray(v, dtype=np.float64)
         for v in (mx[0], mx[1], mx[2], my[0], my[1], mz[0], mz[1], mz[2], c, kappa, rho)]
    lib().ora_ghost(nx, ny, nz, *[_p(v) for v in a], C.c_double(sigma), offp, _p(f))
    return f


def set_kcycle(mask):
    """Levels (bit l) whose coarse correction is two flexible-CG steps
    preconditioned by the cycle below (K-cycle); 0: the V-cycle."""
    lib().ora_set_kcycle(int(mask))


def set_flexible(f):
    """Outer CG beta of AC.acpcg: False Fletcher-Reeves, True flexible."""
    lib().ora_set_flexible(1 if f else 0)


def set_sum_block(b):
    """0: sequential sums (the reference's order); b > 0: blocks of b terms."""
    lib().ora_set_sum_block(int(b))


def spmv(A, x):
    y = np.zeros(A.n)
    lib().ora_spmv(C.byref(A.s), _p(np.ascontiguousarray(x, dtype=np.float64)), _p(y))
    return y


def resid(A, b, x):
    r = np.zeros(A.n)
    lib().ora_resid(C.byref(A.s), _p(np.ascontiguousarray(b, dtype=np.float64)),
                    _p(np.ascontiguousarray(x, dtype=np.float64)), _p(r))
    return r


def zline_solve(A, t):
    """x = T^-1 t with T the tridiagonal (c, -b) part of A along k, per column."""
    x = np.zeros(A.n)
    lib().ora_zline_solve(C.byref(A.s), _p(np.ascontiguousarray(t, dtype=np.float64)), _p(x))
    return x


def dot(x, y):
    return lib().ora_dot(C.c_size_t(x.size), _p(x), _p(y))


def dot_voldiv_avg(x, y, dV):
    return lib().ora_dot_voldiv_avg(C.c_size_t(x.size), _p(x), _p(y), _p(dV))


def _krylov(fn, A, x, b, dV, mini=1, maxit=20, res_rat=0.01, res_tol=1e-15,
            scale=1.0, stalecount=-1):
    x = np.array(x, dtype=np.float64, copy=True)
    hist = np.full(maxit + 1, np.nan)
    R = OraResult()
    conv = fn(C.byref(A.s), _p(x), _p(np.ascontiguousarray(b, dtype=np.float64)),
              _p(np.ascontiguousarray(dV, dtype=np.float64)), mini, maxit,
              C.c_double(res_rat), C.c_double(res_tol), C.c_double(scale), stalecount,
              _p(hist), C.byref(R))
    nh = 1 if R.reason == 7 else min(R.iters, maxit) + 1
    return x, dict(converged=bool(conv), iters=R.iters, reason=R.reason, res0=R.res0,
                   res=R.res, hist=hist[:nh])


def cg(A, x, b, dV, **kw):
    """Reference CG::solve semantics (cg.cpp:10-182)."""
    return _krylov(lib().ora_cg, A, x, b, dV, **kw)


PREC = {"di": 0, "ic0": 1, "ic2": 2, "rbsgs": 10}


def cg_prec(A, prec, x, b, dV, **kw):
    """CG::solve with Prec::di / ic0 / ic2 (preconditioner_form.cpp,
    preconditioner_solve.cpp) or the red-black symmetric Gauss-Seidel
    ("rbsgs") the library runs in place of the incomplete Cholesky."""
    k = PREC[prec]

    def fn(*a):
        return lib().ora_cg_prec(a[0], k, *a[1:])
    return _krylov(fn, A, x, b, dV, **kw)


def prec_apply(A, prec, r):
    z = np.zeros(A.n)
    lib().ora_prec_apply(C.byref(A.s), PREC[prec], _p(np.ascontiguousarray(r, np.float64)),
                         _p(z))
    return z


def bicgs(A, x, b, dV, **kw):
    """Reference BiCGS::solve semantics (bicgs.cpp:9-229)."""
    return _krylov(lib().ora_bicgs, A, x, b, dV, **kw)


def jacobi(A, x, b, mini=1, maxit=20, res_rat=0.01, res_tol=1e-15):
    x = np.array(x, dtype=np.float64, copy=True)
    R = OraResult()
    conv = lib().ora_jacobi(C.byref(A.s), _p(x), _p(b), mini, maxit, C.c_double(res_rat),
                            C.c_double(res_tol), C.byref(R))
    return x, dict(converged=bool(conv), iters=R.iters, reason=R.reason, res0=R.res0, res=R.res)


def levels(rule, nx, ny, nz, maxlev=24):
    lx, ly, lz = (C.c_int * 24)(), (C.c_int * 24)(), (C.c_int * 24)()
    L = lib().ora_levels(rule, nx, ny, nz, maxlev, lx, ly, lz)
    return [(lx[i], ly[i], lz[i]) for i in range(L)]


class AC:
    """Level hierarchy of the additive-correction multigrid."""

    def __init__(self, A, dV, off=None, rule=0, maxlev=24):
        self.ac = OraAC()
        self._keep = (A, dV)
        offp = None
        if off is not None:
            self._off = np.ascontiguousarray(off, dtype=np.uint8).ravel()
            offp = _p(self._off)
        lib().ora_ac_build(C.byref(self.ac), rule, C.byref(A.s), _p(dV), offp, maxlev)

    def __del__(self):
        try:
            lib().ora_ac_free(C.byref(self.ac))
        except Exception:
            pass

    @property
    def nlev(self):
        return self.ac.nlev

    def dims(self, l):
        return (self.ac.nx[l], self.ac.ny[l], self.ac.nz[l])

    def _arr(self, ptr, l):
        n = int(np.prod(self.dims(l)))
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def mat(self, l):
        m = self.ac.A[l]
        return {k: self._arr(getattr(m, k), l).copy() for k in "cwesnbt"}

    def aflag(self, l):
        return self._arr(self.ac.aflag[l], l).copy()

    def phi(self, l):
        return self._arr(self.ac.phi[l], l)

    def f(self, l):
        return self._arr(self.ac.f[l], l)

    def restriction(self, l):
        lib().ora_restriction(C.byref(self.ac), l)

    def interpolation(self, l):
        lib().ora_interpolation(C.byref(self.ac), l)

    def vcycle(self, phi0, f0, toler=-1.0, factor=1e-8, max_cyc=20, min_cyc=0,
               mv=(20, 20, 20), dt=1.0):
        """AC::cycle(none, V, ...) with CG-di smoothers (additive_cycle.cpp)."""
        phi = np.array(phi0, dtype=np.float64, copy=True)
        f0 = np.ascontiguousarray(f0, dtype=np.float64)
        hist = np.full(max_cyc + 1, np.nan)
        ncyc = C.c_int(0)
        mvv = (C.c_int * 3)(*mv)
        ok = lib().ora_ac_vcycle(C.byref(self.ac), _p(phi), _p(f0), C.c_double(toler),
                                 C.c_double(factor), max_cyc, min_cyc, mvv,
                                 C.c_double(dt), _p(hist), C.byref(ncyc))
        return phi, dict(converged=bool(ok), ncyc=ncyc.value, hist=hist[:ncyc.value + 1])

    def cycle(self, phi0, f0, loop="V", init=None, toler=-1.0, factor=1e-8, max_cyc=20,
              min_cyc=0, mv=(20, 20, 20), dt=1.0, stop_if_div=True, priority_min_cyc=True,
              use_linf=False, resrat=1e-18, restol=1e-18, ms=(-1, -1, -1), max_visits=1000,
              solver="cg"):
        """AC::cycle(init, loop, ...) with CG-di (solver "cg") or BiCGS-di
        ("bicgs") level solves (additive_cycle.cpp); loop in "V", "F", "W",
        "flex", init in "V", "F", "W" or None (no full-cycle start).  A flex
        cycle that reaches max_visits level visits ends the run with status
        "bound" (phi: the previous cycle's)."""
        kinds = {"V": 0, "F": 1, "W": 2, "flex": 3, None: -1}
        P = OraCycle(C.c_double(toler), C.c_double(factor), max_cyc, min_cyc,
                     (C.c_int * 3)(*mv), (C.c_int * 3)(*ms), C.c_double(dt), kinds[init],
                     kinds[loop],
                     int(stop_if_div), int(priority_min_cyc), int(use_linf),
                     C.c_double(resrat), C.c_double(restol), max_visits,
                     {"cg": 0, "bicgs": 1}[solver])
        phi = np.array(phi0, dtype=np.float64, copy=True)
        f0 = np.ascontiguousarray(f0, dtype=np.float64)
        hist = np.full(max_cyc + 2, np.nan)
        hlinf = np.full(max_cyc + 2, np.nan)
        nh, ncyc = C.c_int(0), C.c_int(0)
        ok = lib().ora_ac_cycle(C.byref(self.ac), _p(phi), _p(f0), C.byref(P), _p(hist),
                                _p(hlinf), C.byref(nh), C.byref(ncyc))
        return phi, dict(converged=ok == 1, bound=ok == -1, ncyc=ncyc.value,
                         hist=hist[:nh.value], hist_linf=hlinf[:nh.value])

    def coarse_dense(self):
        m = lib().ora_ac_coarse_dense(C.byref(self.ac))
        if m < 0:
            raise ValueError("singular coarsest level of a non-symmetric matrix")
        return np.ctypeslib.as_array(self.ac.G, shape=(m * m,)).reshape(m, m).copy()

    def precond(self, r, nu1=2, nu2=2, omega=0.8, cfac=1.0, smoother=0):
        if not self.ac.G:
            self.coarse_dense()
        z = np.zeros(r.size)
        lib().ora_ac_precond_s(C.byref(self.ac), 0,
                               _p(np.ascontiguousarray(r, dtype=np.float64)), _p(z), nu1, nu2,
                               C.c_double(omega), C.c_double(cfac), smoother)
        return z

    def acpcg(self, x, b, rtol=1e-8, maxit=500, nu1=2, nu2=2, omega=0.8, cfac=1.0,
              smoother=0, orth=False):
        """orth: also return |p_k+1.A p_k| / (|p_k+1|_A |p_k|_A) per step."""
        x = np.array(x, dtype=np.float64, copy=True)
        hist = np.full(maxit + 1, np.nan)
        ob = np.full(maxit, np.nan)
        if orth:
            lib().ora_set_orth_record(_p(ob), maxit)
        it = C.c_int(0)
        try:
            ok = lib().ora_acpcg_s(C.byref(self.ac), _p(x),
                                   _p(np.ascontiguousarray(b, dtype=np.float64)),
                                   C.c_double(rtol), maxit, nu1, nu2, C.c_double(omega),
                                   C.c_double(cfac), smoother, _p(hist), C.byref(it))
        finally:
            if orth:
                lib().ora_set_orth_record(None, 0)
        out = dict(converged=bool(ok), iters=it.value, hist=hist[:it.value + 1])
        if orth:
            out["orth"] = ob[:max(0, it.value - 1)]
        return x, out

    def acbicgs(self, x, b, dV, nu1=2, nu2=2, omega=0.8, cfac=2.0, **kw):
        """BiCGS (bicgs.cpp:9-229 stopping rules, keywords as bicgs) with the
        V-cycle as its preconditioner."""
        if not self.ac.G:
            self.coarse_dense()

        def fn(A, xp, bp, dVp, mini, maxit, rr, rt, sc, stale, hp, R):
            return lib().ora_acbicgs(C.byref(self.ac), xp, bp, dVp, mini, maxit, rr, rt, sc,
                                     stale, nu1, nu2, C.c_double(omega), C.c_double(cfac),
                                     hp, R)
        return _krylov(fn, self._keep[0], x, b, dV, **kw)


def pcg_rel(A, x, b, rtol=1e-8, maxit=1000):
    x = np.array(x, dtype=np.float64, copy=True)
    hist = np.full(maxit + 1, np.nan)
    it = C.c_int(0)
    ok = lib().ora_pcg_rel(C.byref(A.s), _p(x), _p(np.ascontiguousarray(b, dtype=np.float64)),
                           C.c_double(rtol), maxit, _p(hist), C.byref(it))
    return x, dict(converged=bool(ok), iters=it.value, hist=hist[:it.value + 1])


def sphere_case(n, nsph, rmin, rmax, rho_in, rho_out, seed_sph, seed_vel, centre=None,
                ny=None, nz=None, L=1.0):
    """Uniform-grid case: returns (A, ci, dV, b)."""
    ny = ny or n
    nz = nz or n
    h = L / max(n, ny, nz)
    xn = np.array([h * i for i in range(n + 1)])
    yn = np.array([h * i for i in range(ny + 1)])
    zn = np.array([h * i for i in range(nz + 1)])
    mx, my, mz = metrics(xn), metrics(yn), metrics(zn)
    if centre is not None:
        sph = np.array([[centre[0], centre[1], centre[2], rmin]])
    else:
        sph = spheres(seed_sph, nsph, rmin, rmax, (n * h, ny * h, nz * h))
    rho = density(mx[0], my[0], mz[0], sph, rho_in, rho_out)
    A, ci, dV = discretize(mx[1:], my[1:], mz[1:], rho)
    u, v, w = face_velocity(seed_vel, n, ny, nz)
    b = rhs(mx[1], my[1], mz[1], u, v, w)
    return A, ci, dV, b
