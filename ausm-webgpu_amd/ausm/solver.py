"""Host-side mirror of the reference's operator surface for one solve.

The reference drives its WebGPU passes from `main()` (src/main.js): it uploads
the body / far-field rows (`updateGridBoundaries`, src/grid.js:267-296), runs
the grid passes (`prepareGrid`, main.js:658-707), fills the state
(`prepareState`, main.js:631-655), swaps flux pipelines (`updateSolver`,
main.js:629) and steps inside `render()` (main.js:742-792).  `Solver` keeps
those names and their meaning over the C ABI of libausm; every call runs on
the GPU through include/ausm.h.
"""
import ctypes
import math

import numpy as np

from . import lib as L
from .lib import check, f32, f64, ptr
from .lib import load as _load_lib


def naca4_digits(code):
    """NACA 4-digit code -> (m, p, t), as generateNACA4Boundary splits it
    (src/grid.js:41-43)."""
    return (code // 1000) / 100.0, (code // 100) % 10 / 10.0, code % 100 / 100.0


def naca4_boundary(M, code=None, m=0.0, p=0.0, t=0.12, size=0.5, x_offset=0.0, exponent=1.4,
                   spacing="power"):
    """Body row of M points, CCW from the trailing edge (src/grid.js:38-52).
    Chord = 2*size centred at x_offset; spacing "power" is the reference's
    x = (1-s)^exponent law, "cosine" the usual cosine clustering."""
    if code is not None:
        m, p, t = naca4_digits(code)
    out = np.zeros((M, 2), dtype=np.float64)
    sp = L.SPACING_COSINE if spacing == "cosine" else L.SPACING_POWER
    check(L.load().ausm_boundary_naca4(m, p, t, size, x_offset, exponent, sp, M, ptr(out)))
    return out


def cst_boundary(M, a_upper, a_lower, te_gap=0.0, size=0.5, x_offset=0.0, exponent=1.4,
                 spacing="power"):
    """Kulfan class-shape-transformation body of M points, in the point order
    and placement of naca4_boundary (no counterpart in the reference, which
    draws NACA 4-digit bodies only).  a_lower is signed (negative below the
    chord); te_gap = 0 closes the trailing edge."""
    au, al = f64(a_upper), f64(a_lower)
    if au.shape != al.shape or au.ndim != 1:
        raise ValueError("a_upper and a_lower need the same number of coefficients")
    out = np.zeros((M, 2), dtype=np.float64)
    sp = L.SPACING_COSINE if spacing == "cosine" else L.SPACING_POWER
    check(L.load().ausm_boundary_cst(ptr(au), ptr(al), au.size, te_gap, size, x_offset, exponent,
                                     sp, M, ptr(out)))
    return out


def bezier_boundary(M, upper, lower, size=0.5, x_offset=0.0, exponent=1.4, spacing="power"):
    """Bezier body of M points (no counterpart in the reference): upper and
    lower are (n, 2) control polygons in chord units from the shared
    leading-edge point (t = 0) to each trailing-edge point (t = 1); the curve
    parameter takes naca4_boundary's chordwise stations and point order."""
    cu = np.ascontiguousarray(np.asarray(upper, dtype=np.float64).reshape(-1, 2))
    cl = np.ascontiguousarray(np.asarray(lower, dtype=np.float64).reshape(-1, 2))
    out = np.zeros((M, 2), dtype=np.float64)
    sp = L.SPACING_COSINE if spacing == "cosine" else L.SPACING_POWER
    check(L.load().ausm_boundary_bezier(ptr(cu), cu.shape[0], ptr(cl), cl.shape[0], size, x_offset,
                                        exponent, sp, M, ptr(out)))
    return out


def sears_haack_boundary(M, V=0.0002, L=0.4, x_offset=-0.7):
    """Sears-Haack body of volume V and length L (src/grid.js:53-61; defaults
    are the reference's object menu, src/grid.js:262)."""
    out = np.zeros((M, 2), dtype=np.float64)
    check(_load_lib().ausm_boundary_sears_haack(V, L, x_offset, M, ptr(out)))
    return out


def haack_projectile_boundary(M, F=6.0, R=0.03, L=0.5, Rb=0.5, Lb=0.2, C=0.0, x_offset=-0.7):
    """Haack-series nose cone on a body with a boat tail (src/grid.js:75-118)."""
    out = np.zeros((M, 2), dtype=np.float64)
    check(_load_lib().ausm_boundary_haack_projectile(F, R, L, Rb, Lb, C, x_offset, M, ptr(out)))
    return out


def polygon_boundary(M, R=0.1, N=4, F=0.1, a=0.0, x_offset=-0.7):
    """Regular N-gon body (src/grid.js:224-231); the reference's circle body
    is N = 400, F = 1 (src/grid.js:258-259)."""
    out = np.zeros((M, 2), dtype=np.float64)
    check(L.load().ausm_boundary_polygon(R, N, F, a, x_offset, M, ptr(out)))
    return out


def dat_boundary(M, text, size=0.3, x_offset=-0.7):
    """Airfoil coordinate file (Selig or Lednicer) resampled to M points
    (src/airfoilDatParser.js); returns (points, format)."""
    if isinstance(text, str):
        text = text.encode()
    out = np.zeros((M, 2), dtype=np.float64)
    fmt = ctypes.c_int(-1)
    check(L.load().ausm_boundary_dat(text, size, x_offset, M, ptr(out), ctypes.byref(fmt)))
    return out, ("selig", "lednicer")[fmt.value]


def circle_boundary(M, radius):
    """Circular far field (src/grid.js:232-237)."""
    out = np.zeros((M, 2), dtype=np.float64)
    check(L.load().ausm_boundary_circle(radius, M, ptr(out)))
    return out


def rect_boundary(M, ratio=1.5):
    """The reference's default rectangular far field (src/grid.js:238-255, 265)."""
    out = np.zeros((M, 2), dtype=np.float64)
    check(L.load().ausm_boundary_rect(ratio, M, ptr(out)))
    return out


def freestream(mach, aoa_deg=0.0, rho=1.0, p=None, gamma=1.4):
    """Free-stream (rho, p, vx, vy, gamma) for a Mach number and angle of
    attack; with the reference defaults rho = 1, p = 1/gamma the speed of
    sound is 1 and the velocity equals the Mach number (src/util.js:92-94)."""
    if p is None:
        p = 1.0 / gamma
    a = math.sqrt(gamma * p / rho)
    v = mach * a
    th = math.radians(aoa_deg)
    return rho, p, v * math.cos(th), v * math.sin(th), gamma


class Solver:
    """One solve of M x N cells resident on one GPU."""

    def __init__(self, M, N, device=0):
        self._L = L.load()
        h = ctypes.c_void_p()
        check(self._L.ausm_ctx_create(device, M, N, ctypes.byref(h)))
        self.h = h
        self.M, self.N = M, N
        self._fs = None

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if self.h:
            self._L.ausm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        s = ctypes.c_void_p()
        check(self._L.ausm_ctx_stream(self.h, ctypes.byref(s)))
        return s.value

    def sync(self):
        check(self._L.ausm_sync(self.h))

    # -- grid (updateGridBoundaries + prepareGrid) -------------------------
    def update_grid_boundaries(self, body_xy, outer_xy):
        self._body = f64(body_xy).reshape(self.M, 2)
        self._outer = f64(outer_xy).reshape(self.M, 2)

    def prepare_grid(self, iters=10000, method="jacobi", omega=1.0):
        meth = L.GRID_SOR if method == "sor" else L.GRID_JACOBI
        check(self._L.ausm_grid_from_boundary(self.h, ptr(self._body), ptr(self._outer), iters,
                                              meth, omega))

    def set_vertices(self, xy):
        """Vertex upload; float64 input keeps full precision for the FP64
        solve (the FP32 solve always uses the vertices rounded to FP32)."""
        if np.asarray(xy).dtype == np.float64:
            a = f64(xy).reshape(self.N + 1, self.M, 2)
            check(self._L.ausm_grid_set_vertices64(self.h, ptr(a)))
        else:
            a = f32(xy).reshape(self.N + 1, self.M, 2)
            check(self._L.ausm_grid_set_vertices(self.h, ptr(a)))

    def vertices(self, fp64=False):
        out = np.zeros((self.N + 1, self.M, 2), dtype=np.float64 if fp64 else np.float32)
        fn = self._L.ausm_grid_get_vertices64 if fp64 else self._L.ausm_grid_get_vertices
        check(fn(self.h, ptr(out)))
        return out

    def metrics(self):
        area = np.zeros((self.N, self.M), dtype=np.float32)
        flen = np.zeros((self.N, self.M, 4), dtype=np.float32)
        check(self._L.ausm_grid_get_metrics(self.h, ptr(area), ptr(flen)))
        return area, flen

    # -- uniforms --------------------------------------------------------
    def set_freestream(self, rho, p, vx, vy, gamma=1.4):
        check(self._L.ausm_set_freestream(self.h, rho, p, vx, vy, gamma))
        self._fs = (rho, p, vx, vy, gamma)

    def set_flow(self, mach, aoa_deg=0.0, **kw):
        self.set_freestream(*freestream(mach, aoa_deg, **kw))

    def set_scheme(self, scheme=L.SLAU2, limiter=L.VANALBADA, K_p=0.0, K_u=0.75):
        if isinstance(scheme, str):
            scheme = L.SCHEMES[scheme]
        if isinstance(limiter, str):
            limiter = L.LIMITERS[limiter]
        self._scheme = (scheme, limiter, K_p, K_u)
        check(self._L.ausm_set_scheme(self.h, scheme, limiter, K_p, K_u))

    def update_solver(self, name):
        """Swap the flux scheme, keeping the state (main.js:629)."""
        s = getattr(self, "_scheme", (L.SLAU2, L.VANALBADA, 0.0, 0.75))
        self.set_scheme(L.SCHEMES[name], s[1], s[2], s[3])

    def set_cfl(self, cfl=1.5, maxdt=1e-4, steps_per_dt_refresh=1):
        check(self._L.ausm_set_cfl(self.h, cfl, maxdt, steps_per_dt_refresh))

    def set_math(self, mode):
        if isinstance(mode, str):
            mode = L.MATH_FAST if mode == "fast" else L.MATH_EXACT
        check(self._L.ausm_set_math(self.h, mode))

    def set_tile(self, rows):
        check(self._L.ausm_set_tile(self.h, rows))

    def set_dt_mode(self, mode):
        """"global" (the reference's single CFL step) or "local" (each cell
        its own CFL step from Q^n: a steady-state accelerator)."""
        if isinstance(mode, str):
            mode = {"global": L.DT_GLOBAL, "local": L.DT_LOCAL}[mode]
        check(self._L.ausm_set_dt_mode(self.h, mode))

    def freeze_limiter(self, on=True):
        """Limiter freezing (van Albada limiters; no counterpart in the
        reference): record psi = phi(r)/r of every face from the current
        state and reconstruct with it from now on, until freeze_limiter(False)
        or a new state, grid, limiter or precision.  A convergence device for
        MUSCL solves whose limiter switching holds the residual in a limit
        cycle."""
        check(self._L.ausm_freeze_limiter(self.h, 1 if on else 0))

    @property
    def limiter_frozen(self):
        on = ctypes.c_int()
        check(self._L.ausm_limiter_frozen(self.h, ctypes.byref(on)))
        return bool(on.value)

    def set_precision(self, prec):
        """"fp32" (the reference's arithmetic), "fp64" (the comparison solve)
        or "fp32c" (FP32 arithmetic, compensated hi + lo state accumulation);
        the current solution is carried over."""
        if isinstance(prec, str):
            prec = {"fp32": L.PREC_FP32, "fp64": L.PREC_FP64, "fp32c": L.PREC_FP32C}[prec]
        check(self._L.ausm_set_precision(self.h, prec))
        self._prec = prec

    @property
    def fp64(self):
        return getattr(self, "_prec", L.PREC_FP32) == L.PREC_FP64

    # -- state and stepping --------------------------------------------------
    def prepare_state(self, freestream=True):
        """Fill the state (main.js:631-655); the reference starts from rest,
        freestream=True starts from the free stream instead."""
        check(self._L.ausm_init_state(self.h, L.INIT_FREESTREAM if freestream else L.INIT_REST))

    def set_state(self, q):
        if self.fp64:
            a = f64(q).reshape(self.N + 3, self.M, 4)
            check(self._L.ausm_set_state64(self.h, ptr(a)))
        else:
            a = f32(q).reshape(self.N + 3, self.M, 4)
            check(self._L.ausm_set_state(self.h, ptr(a)))

    def set_state64(self, q):
        """FP64 / FP32C precision: set the state from float64 (FP32C: split
        into hi + lo)."""
        a = f64(q).reshape(self.N + 3, self.M, 4)
        check(self._L.ausm_set_state64(self.h, ptr(a)))

    def state64(self):
        """FP64 / FP32C precision: the state as float64 (FP32C: hi + lo)."""
        out = np.zeros((self.N + 3, self.M, 4), dtype=np.float64)
        check(self._L.ausm_get_state64(self.h, ptr(out)))
        return out

    def state(self):
        """(N+3) x M x 4 in the reference's texture layout; float64 in the
        FP64 precision."""
        if self.fp64:
            out = np.zeros((self.N + 3, self.M, 4), dtype=np.float64)
            check(self._L.ausm_get_state64(self.h, ptr(out)))
        else:
            out = np.zeros((self.N + 3, self.M, 4), dtype=np.float32)
            check(self._L.ausm_get_state(self.h, ptr(out)))
        return out

    def prepare_steps(self, n):
        """Capture the step graphs a following step(n) launches, running
        nothing (graph set-up kept out of a timed call)."""
        check(self._L.ausm_prepare_steps(self.h, n))

    def step(self, n=1, info=True):
        """Advance n TVD-RK3 steps (the render() step loop, main.js:744-777)."""
        if not info:
            check(self._L.ausm_step(self.h, n, None))
            return None
        si = L.StepInfo()
        check(self._L.ausm_step(self.h, n, ctypes.byref(si)))
        return {"dt": si.dt, "max_wave_speed": si.max_wave_speed, "res_l2": si.res_l2,
                "nonfinite": bool(si.nonfinite), "steps_done": si.steps_done}

    def residual(self):
        if self.fp64:
            out = np.zeros((self.N, self.M, 4), dtype=np.float64)
            check(self._L.ausm_eval_residual64(self.h, ptr(out)))
        else:
            out = np.zeros((self.N, self.M, 4), dtype=np.float32)
            check(self._L.ausm_eval_residual(self.h, ptr(out)))
        return out

    def wave_speeds(self):
        ws = np.zeros((self.N, self.M), dtype=np.float32)
        mx = ctypes.c_float()
        check(self._L.ausm_wave_speeds(self.h, ptr(ws), ctypes.byref(mx)))
        return ws, mx.value

    def forces(self, ref_length=1.0, want_cp=False):
        out = np.zeros(4, dtype=np.float64)
        cp = np.zeros(self.M, dtype=np.float64)
        check(self._L.ausm_forces(self.h, ref_length, ptr(out), ptr(cp)))
        res = {"fx": out[0], "fy": out[1], "cl": out[2], "cd": out[3]}
        if want_cp:
            res["cp"] = cp
        return res

    def derived_field(self, kind):
        """Display field of the current solution (the visualization pass,
        ausm.js:586-710, before its colour map): kind is a name of
        lib.FIELDS or its number.  Returns (N x M float32, free-stream value)."""
        if isinstance(kind, str):
            kind = L.FIELDS[kind]
        out = np.zeros((self.N, self.M), dtype=np.float32)
        fs = ctypes.c_float()
        check(self._L.ausm_derived_field(self.h, kind, ptr(out), ctypes.byref(fs)))
        return out, fs.value

    def residual_norm(self):
        v = ctypes.c_double()
        check(self._L.ausm_residual_norm(self.h, ctypes.byref(v)))
        return v.value
