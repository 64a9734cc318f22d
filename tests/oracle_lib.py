"""ctypes bindings for the CPU oracle (oracle/liborc.so) — test infrastructure.

The oracle is a plain-C restatement of the reference solver (ref32 = FP32,
ref64 = FP64).  Tests use it as the checker for the HIP library; nothing in
the product imports this module.
"""
import contextlib
import ctypes
import math
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIBORC = os.path.join(ORACLE_DIR, "liborc.so")

SLAU2, SLAU, AUSMUP = 0, 1, 2
VANALBADA, MINMOD, VANLEER, NONE, VANALBADA_CLAMPED = 0, 1, 2, 3, 4
SCHEMES = {"SLAU2": SLAU2, "SLAU": SLAU, "AUSM+-up": AUSMUP}

P_GAMMA, P_INRHO, P_INP, P_VX, P_VY = 0, 1, 2, 3, 4
P_IS0, P_MAXDT, P_CFL, P_KP, P_KU = 5, 9, 10, 11, 12
P_COUNT = 13


LIBBASE = os.path.join(ORACLE_DIR, "libbaseline.so")


def ensure_built():
    if not os.path.exists(LIBORC) or not os.path.exists(LIBBASE):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    return LIBORC


_lib = None
_libomp = None


def libomp():
    """ref32 built -O3 with OpenMP over rows (oracle/libbaseline.so, prefix
    base32_): the same FP32 code without contraction, for long runs."""
    global _libomp
    if _libomp is None:
        ensure_built()
        _libomp = ctypes.CDLL(LIBBASE)
        _declare(_libomp, (("base32_", ctypes.c_float),), geometry=False)
    return _libomp


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(ensure_built())
        _declare(_lib)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _declare(L, prefixes=(("orc32_", ctypes.c_float), ("orc64_", ctypes.c_double)),
             geometry=True):
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    for pfx, cr in prefixes:
        g = lambda n: getattr(L, pfx + n)
        g("face_flux").argtypes = [vp, ci, ci, vp, vp, vp, vp, cr, cr, vp]
        g("flux_pass").argtypes = [vp, ci, ci, ci, ci, ci, vp, vp, vp]
        g("boundary_pass").argtypes = [vp, ci, ci, vp, vp, vp]
        g("residual_pass").argtypes = [ci, ci, vp, vp, vp, vp, vp]
        g("dt_from_ws").argtypes = [vp, cr]
        g("dt_from_ws").restype = cr
        g("stage_pass").argtypes = [ci, ci, ci, cr, vp, vp, vp, vp]
        g("max_wave_speed").argtypes = [vp, ci, ci, vp, vp, vp, vp]
        g("max_wave_speed").restype = cr
        g("grid_finalize").argtypes = [ci, ci, vp, vp, vp]
        g("grid_interpolate").argtypes = [ci, ci, vp]
        g("grid_jacobi_sweep").argtypes = [ci, ci, vp, vp]
        g("grid_jacobi").argtypes = [ci, ci, vp, ci]
        g("forces").argtypes = [vp, ci, vp, vp, cd, vp, vp]
        g("init_state").argtypes = [vp, ci, ci, ci, vp]
        g("eval_residual").argtypes = [vp, ci, ci, ci, ci, vp, vp, vp, vp, vp]
        g("run_steps").argtypes = [vp, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, vp, vp, vp]
        g("run_steps").restype = cr
        g("run_steps_dt").argtypes = [vp, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, vp, vp,
                                      vp, ci]
        g("run_steps_dt").restype = cr
        g("run_steps_comp").argtypes = [vp, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, ci,
                                        vp, vp, ci]
        g("run_steps_comp").restype = cr
        g("derived_field").argtypes = [vp, ci, ci, vp, ci, vp, vp]
        g("residual_norm").argtypes = [ci, ci, vp, ci]
        g("residual_norm").restype = cd
        g("set_freeze").argtypes = [ci, vp]
    if not geometry:
        return
    for n in ("geo_naca4_point",):
        getattr(L, n).argtypes = [cd, cd, cd, cd, cd, cd, ci, ci, ci, vp]
    L.geo_circle_point.argtypes = [cd, ci, ci, vp]
    L.geo_cst_point.argtypes = [vp, vp, ci, cd, cd, cd, cd, ci, ci, ci, vp]
    L.geo_bezier_point.argtypes = [vp, ci, vp, ci, cd, cd, cd, ci, ci, ci, vp]
    L.geo_sears_haack_point.argtypes = [cd, cd, cd, ci, ci, vp]
    L.geo_haack_point.argtypes = [cd, cd, cd, cd, cd, cd, cd, ci, ci, vp]
    L.geo_polygon_point.argtypes = [cd, ci, cd, cd, cd, ci, ci, vp]
    L.geo_rect_point.argtypes = [cd, ci, ci, vp]


# ----------------------------------------------------------------------------
# host-side parameter setup (mirrors util.js:76-99 and main.js:778-784)
# ----------------------------------------------------------------------------

def make_params(mach, aoa_deg, gamma=1.4, rho=1.0, p=None, cfl=0.5, maxdt=1e-4,
                K_p=0.0, K_u=0.75, dtype=np.float32):
    """Freestream + solver parameters.  As in the reference, every value is
    computed in double on the host, then stored as FP32 (the uniform buffer
    is f32); the FP64 oracle gets the same FP32-rounded values."""
    if p is None:
        p = 1.0 / gamma
    a = math.sqrt(gamma * p / rho)
    v = mach * a
    aoa = math.radians(aoa_deg)
    vx, vy = v * math.cos(aoa), v * math.sin(aoa)
    rhoE = p / (gamma - 1.0) + 0.5 * (v * v) * rho
    par = np.zeros(P_COUNT, dtype=np.float64)
    par[P_GAMMA], par[P_INRHO], par[P_INP] = gamma, rho, p
    par[P_VX], par[P_VY] = vx, vy
    par[P_IS0:P_IS0 + 4] = [rho, vx * rho, vy * rho, rhoE]
    par[P_MAXDT], par[P_CFL], par[P_KP], par[P_KU] = maxdt, cfl, K_p, K_u
    return par.astype(np.float32).astype(dtype)


def make_params64(fs, cfl=0.5, maxdt=1e-4, K_p=0.0, K_u=0.75):
    """FP64 parameter vector of the all-FP64 solve for the free stream
    fs = (rho, p, vx, vy, gamma) as given (unrounded).  inState is formed in
    double in the order rho, vx*rho, vy*rho, p/(gamma-1) + 0.5*(vx^2+vy^2)*rho;
    the solver settings the C ABI passes as FP32 (cfl, maxdt, K_p, K_u) are
    FP32 values widened."""
    rho, p, vx, vy, gamma = (float(x) for x in fs)
    par = np.zeros(P_COUNT, dtype=np.float64)
    par[P_GAMMA], par[P_INRHO], par[P_INP] = gamma, rho, p
    par[P_VX], par[P_VY] = vx, vy
    par[P_IS0:P_IS0 + 4] = [rho, vx * rho, vy * rho,
                            p / (gamma - 1.0) + 0.5 * (vx * vx + vy * vy) * rho]
    f = np.float32
    par[P_MAXDT], par[P_CFL] = float(f(maxdt)), float(f(cfl))
    par[P_KP], par[P_KU] = float(f(K_p)), float(f(K_u))
    return par


# ----------------------------------------------------------------------------
# geometry
# ----------------------------------------------------------------------------

def naca4_boundary(M, m, p, t, size=0.5, x_offset=0.0, expo=1.4, spacing=0):
    out = np.zeros((M, 2), dtype=np.float64)
    L = lib()
    for i in range(M):
        L.geo_naca4_point(m, p, t, size, x_offset, expo, spacing, M, i, _ptr(out[i]))
    return out


def cst_boundary(M, a_upper, a_lower, te_gap=0.0, size=0.5, x_offset=0.0, expo=1.4, spacing=1):
    au = np.ascontiguousarray(a_upper, dtype=np.float64)
    al = np.ascontiguousarray(a_lower, dtype=np.float64)
    assert au.shape == al.shape
    out = np.zeros((M, 2), dtype=np.float64)
    for i in range(M):
        lib().geo_cst_point(_ptr(au), _ptr(al), au.size, te_gap, size, x_offset, expo, spacing, M,
                            i, _ptr(out[i]))
    return out


def bezier_boundary(M, upper, lower, size=0.5, x_offset=0.0, expo=1.4, spacing=1):
    cu = np.ascontiguousarray(np.asarray(upper, dtype=np.float64).reshape(-1, 2))
    cl = np.ascontiguousarray(np.asarray(lower, dtype=np.float64).reshape(-1, 2))
    out = np.zeros((M, 2), dtype=np.float64)
    for i in range(M):
        lib().geo_bezier_point(_ptr(cu), cu.shape[0], _ptr(cl), cl.shape[0], size, x_offset, expo,
                               spacing, M, i, _ptr(out[i]))
    return out


def _points(M, fn, *args):
    out = np.zeros((M, 2), dtype=np.float64)
    for i in range(M):
        fn(*args, M, i, _ptr(out[i]))
    return out


def sears_haack_boundary(M, V, L, x_offset):
    return _points(M, lib().geo_sears_haack_point, V, L, x_offset)


def haack_boundary(M, F=6.0, R=0.03, L=0.5, Rb=0.5, Lb=0.2, C=0.0, x_offset=-0.7):
    return _points(M, lib().geo_haack_point, F, R, L, Rb, Lb, C, x_offset)


def polygon_boundary(M, R, N, F, a, x_offset):
    return _points(M, lib().geo_polygon_point, R, N, F, a, x_offset)


def naca4_digits(code):
    return (code // 1000) / 100.0, (code // 100) % 10 / 10.0, code % 100 / 100.0


def circle_boundary(M, radius):
    out = np.zeros((M, 2), dtype=np.float64)
    for i in range(M):
        lib().geo_circle_point(radius, M, i, _ptr(out[i]))
    return out


def rect_boundary(M, ratio):
    out = np.zeros((M, 2), dtype=np.float64)
    for i in range(M):
        lib().geo_rect_point(ratio, M, i, _ptr(out[i]))
    return out


class Oracle:
    """One precision of the oracle (32 -> ref32, 64 -> ref64)."""

    def __init__(self, prec=32, omp=False):
        """omp: ref32 from the OpenMP build (bitwise the same results)."""
        self.prec = prec
        self.dt = np.float32 if prec == 32 else np.float64
        if omp:
            assert prec == 32
            self.L = libomp()
            self.pfx = "base32_"
        else:
            self.L = lib()
            self.pfx = "orc32_" if prec == 32 else "orc64_"

    def f(self, name):
        return getattr(self.L, self.pfx + name)

    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.dt)

    @contextlib.contextmanager
    def freeze(self, mode, psi):
        """Limiter freezing for the flux passes inside the block: mode 1
        records psi into `psi`, mode 2 reconstructs with it.  psi: a flat
        array of (2N+1) x M x 12 values of this precision."""
        assert psi.dtype == self.dt and psi.flags.c_contiguous
        self.f("set_freeze")(mode, _ptr(psi))
        try:
            yield
        finally:
            self.f("set_freeze")(0, None)

    def psi_array(self, M, N):
        return np.zeros((2 * N + 1) * M * 12, dtype=self.dt)

    # grid -------------------------------------------------------------
    def grid(self, body, outer, N, iters):
        """Interpolation + Jacobi TTM, in this precision.  Returns (N+1,M,2)."""
        M = body.shape[0]
        v = np.zeros((N + 1, M, 2), dtype=self.dt)
        v[0] = body
        v[N] = outer
        self.f("grid_interpolate")(M, N, _ptr(v))
        if iters:
            self.f("grid_jacobi")(M, N, _ptr(v), iters)
        return v

    def jacobi_sweep(self, vtx):
        N1, M, _ = vtx.shape
        vin = self.arr(vtx)
        out = np.zeros_like(vin)
        self.f("grid_jacobi_sweep")(M, N1 - 1, _ptr(vin), _ptr(out))
        return out

    def finalize(self, vtx):
        vtx = self.arr(vtx)
        N1, M, _ = vtx.shape
        N = N1 - 1
        area = np.zeros((N, M), dtype=self.dt)
        flen = np.zeros((N, M, 4), dtype=self.dt)
        self.f("grid_finalize")(M, N, _ptr(vtx), _ptr(area), _ptr(flen))
        return area, flen

    # solver passes ----------------------------------------------------
    def init_state(self, par, M, N, freestream=True):
        st = np.zeros((N + 3, M, 4), dtype=self.dt)
        self.f("init_state")(_ptr(self.arr(par)), M, N, int(freestream), _ptr(st))
        return st

    def boundary(self, par, vtx, st):
        vtx, st = self.arr(vtx), self.arr(st)
        out = np.zeros_like(st)
        N = vtx.shape[0] - 1
        M = vtx.shape[1]
        self.f("boundary_pass")(_ptr(self.arr(par)), M, N, _ptr(vtx), _ptr(st), _ptr(out))
        return out

    def flux(self, par, scheme, limiter, vertical, vtx, st):
        vtx, st = self.arr(vtx), self.arr(st)
        N = vtx.shape[0] - 1
        M = vtx.shape[1]
        shape = (N + 1, M, 4) if vertical else (N, M + 1, 4)
        out = np.zeros(shape, dtype=self.dt)
        self.f("flux_pass")(_ptr(self.arr(par)), scheme, limiter, int(vertical), M, N, _ptr(vtx),
                            _ptr(st), _ptr(out))
        return out

    def face_flux(self, par, scheme, limiter, QL2, QL, QR, QR2, n):
        F = np.zeros(4, dtype=self.dt)
        q = [self.arr(x) for x in (QL2, QL, QR, QR2)]
        self.f("face_flux")(_ptr(self.arr(par)), scheme, limiter, *[_ptr(x) for x in q],
                            float(n[0]), float(n[1]), _ptr(F))
        return F

    def residual(self, par, scheme, limiter, vtx, st, area=None, flen=None):
        vtx, st = self.arr(vtx), self.arr(st)
        if area is None:
            area, flen = self.finalize(vtx)
        N = vtx.shape[0] - 1
        M = vtx.shape[1]
        res = np.zeros((N, M, 4), dtype=self.dt)
        self.f("eval_residual")(_ptr(self.arr(par)), scheme, limiter, M, N, _ptr(vtx),
                                _ptr(self.arr(flen)), _ptr(self.arr(area)), _ptr(st), _ptr(res))
        return res

    def max_ws(self, par, st, area, flen, want_field=False):
        N, M = area.shape
        ws = np.zeros((N, M), dtype=self.dt) if want_field else None
        v = self.f("max_wave_speed")(_ptr(self.arr(par)), M, N, _ptr(self.arr(st)),
                                     _ptr(self.arr(flen)), _ptr(self.arr(area)),
                                     _ptr(ws) if want_field else None)
        return (v, ws) if want_field else v

    def derived_field(self, par, st, mode):
        """(N x M field, free-stream value) of the visualization pass."""
        st = self.arr(st)
        N = st.shape[0] - 3
        M = st.shape[1]
        out = np.zeros((N, M), dtype=self.dt)
        fs = np.zeros(1, dtype=self.dt)
        self.f("derived_field")(_ptr(self.arr(par)), M, N, _ptr(st), mode, _ptr(out), _ptr(fs))
        return out, fs[0]

    def forces(self, par, st, vtx, ref_len=1.0, want_cp=False):
        M = vtx.shape[1]
        out = np.zeros(4, dtype=np.float64)
        cp = np.zeros(M, dtype=np.float64)
        self.f("forces")(_ptr(self.arr(par)), M, _ptr(self.arr(st)), _ptr(self.arr(vtx)),
                         float(ref_len), _ptr(out), _ptr(cp))
        return (out, cp) if want_cp else out


class Solver:
    """The reference's three-texture step driver on the oracle."""

    def __init__(self, prec, par, scheme, limiter, vtx, refresh=1, freestream=True,
                 vtx_metric=None, local_dt=False, comp=False, omp=False):
        """comp: compensated-state stages (solution = state + lo); omp: the
        OpenMP build of ref32."""
        self.o = Oracle(prec, omp=omp)
        self.local_dt = int(bool(local_dt))
        self.comp = bool(comp)
        self.par = self.o.arr(par)
        self.scheme, self.limiter = scheme, limiter
        self.vtx = self.o.arr(vtx)
        vm = self.vtx if vtx_metric is None else self.o.arr(vtx_metric)
        self.area, self.flen = self.o.finalize(vm)
        self.N = self.vtx.shape[0] - 1
        self.M = self.vtx.shape[1]
        self.refresh = refresh
        s = self.o.init_state(self.par, self.M, self.N, freestream)
        self.s0, self.s1, self.s2 = s.copy(), s.copy(), s.copy()
        self.count = ctypes.c_longlong(0)
        self.maxws = np.zeros(1, dtype=self.o.dt)
        self.maxws[0] = self.o.max_ws(self.par, self.s2, self.area, self.flen)
        self.last_dt = None
        self.lo = np.zeros_like(self.s2)
        self.psi = None  # recorded limiter (freeze_limiter)

    def set_state(self, st):
        st = self.o.arr(st)
        self.s0[...] = st
        self.s1[...] = st
        self.s2[...] = st
        self.lo[...] = 0
        self.psi = None  # a new state ends limiter freezing
        self.maxws[0] = self.o.max_ws(self.par, self.s2, self.area, self.flen)

    def freeze_limiter(self, on=True):
        """Record psi from the current state (a residual pass) and step with
        the frozen reconstruction from now on; on=False ends it."""
        if not on:
            self.psi = None
            return
        psi = self.o.psi_array(self.M, self.N)
        with self.o.freeze(1, psi):
            self.o.residual(self.par, self.scheme, self.limiter, self.vtx, self.s2, self.area,
                            self.flen)
        self.psi = psi

    def residual(self):
        """L(Q) of the current state (frozen if the limiter is)."""
        with self._frozen():
            return self.o.residual(self.par, self.scheme, self.limiter, self.vtx, self.s2,
                                   self.area, self.flen)

    def _frozen(self):
        if self.psi is None:
            return contextlib.nullcontext()
        return self.o.freeze(2, self.psi)

    def step(self, n, want_res=False):
        with self._frozen():
            return self._step(n, want_res)

    def _step(self, n, want_res=False):
        if self.comp:
            assert not want_res
            self.last_dt = self.o.f("run_steps_comp")(
                _ptr(self.par), self.scheme, self.limiter, self.M, self.N, _ptr(self.vtx),
                _ptr(self.flen), _ptr(self.area), _ptr(self.s0), _ptr(self.s1), _ptr(self.s2),
                _ptr(self.lo), n, self.refresh, ctypes.byref(self.count), _ptr(self.maxws),
                self.local_dt)
            return None
        res = np.zeros((self.N, self.M, 4), dtype=self.o.dt) if want_res else None
        dt = self.o.f("run_steps_dt")(_ptr(self.par), self.scheme, self.limiter, self.M, self.N,
                                      _ptr(self.vtx), _ptr(self.flen), _ptr(self.area),
                                      _ptr(self.s0), _ptr(self.s1), _ptr(self.s2), n,
                                      self.refresh, ctypes.byref(self.count), _ptr(self.maxws),
                                      _ptr(res) if want_res else None, self.local_dt)
        self.last_dt = dt
        return res

    @property
    def state(self):
        return self.s2
