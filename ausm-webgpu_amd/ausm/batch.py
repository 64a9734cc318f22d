"""Batched solves over the C ABI (ausm_batch_*).

A `Batch` keeps many independent solves of different sizes resident in HBM
and advances all running ones together (one persistent kernel per RK stage
walking a tile list over every grid).  Requests are queued and admitted as
table slots and arena memory free up (continuous batching).
"""
import ctypes

import numpy as np

from . import lib as L
from .lib import check, f32, ptr

STATES = {0: "queued", 1: "running", 2: "converged", 3: "capped", 4: "failed"}


class Batch:
    def __init__(self, max_solves=64, hbm_budget_bytes=0, device=0, math="fast",
                 precision="fp32"):
        self._L = L.load()
        h = ctypes.c_void_p()
        check(self._L.ausm_batch_create(device, max_solves, hbm_budget_bytes, ctypes.byref(h)))
        self.h = h
        self.set_math(math)
        self.set_precision(precision)
        self._keep = {}

    def set_precision(self, prec):
        """"fp32" or "fp32c" (compensated state) for every solve; before the
        first submit."""
        if isinstance(prec, str):
            prec = {"fp32": L.PREC_FP32, "fp32c": L.PREC_FP32C}[prec]
        check(self._L.ausm_batch_set_precision(self.h, prec))

    def close(self):
        if self.h:
            self._L.ausm_batch_destroy(self.h)
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
        check(self._L.ausm_batch_stream(self.h, ctypes.byref(s)))
        return s.value

    def set_math(self, mode):
        check(self._L.ausm_batch_set_math(self.h, L.MATH_FAST if mode == "fast" else L.MATH_EXACT))

    def set_schedule(self, rows_per_wave=0, blocks_per_cu=0):
        check(self._L.ausm_batch_set_schedule(self.h, rows_per_wave, blocks_per_cu))

    def submit(self, vertices, freestream, scheme=L.SLAU2, limiter=L.VANALBADA, K_p=0.0,
               K_u=0.75, cfl=0.5, maxdt=1.0, init="freestream", max_steps=1000, res_drop=0.0,
               ref_length=1.0, dt_mode="global", freeze_after=0):
        """Queue one solve.  vertices: (N+1, M, 2) FP32; freestream: (rho, p,
        vx, vy, gamma); dt_mode "global" (the reference's) or "local" (each
        cell its own CFL step); freeze_after > 0 freezes a van Albada limiter
        (Solver.freeze_limiter) at the first convergence check at or after
        that many steps.  Returns the solve id."""
        v = f32(vertices)
        N1, M, _ = v.shape
        rho, p, vx, vy, gamma = freestream
        if isinstance(scheme, str):
            scheme = L.SCHEMES[scheme]
        d = L.SolveDesc(M=M, N=N1 - 1, vertices=ptr(v), rho=rho, p=p, vx=vx, vy=vy, gamma=gamma,
                        scheme=scheme, limiter=limiter, K_p=K_p, K_u=K_u, cfl=cfl, maxdt=maxdt,
                        init_mode=L.INIT_FREESTREAM if init == "freestream" else L.INIT_REST,
                        max_steps=max_steps, res_drop=res_drop, ref_length=ref_length,
                        dt_mode=L.DT_LOCAL if dt_mode == "local" else L.DT_GLOBAL,
                        freeze_after=freeze_after)
        sid = ctypes.c_uint64()
        check(self._L.ausm_batch_submit(self.h, ctypes.byref(d), ctypes.byref(sid)))
        self._keep[sid.value] = (M, N1 - 1)
        return sid.value

    def advance(self, nsteps, check_every=50):
        pend = ctypes.c_int()
        check(self._L.ausm_batch_advance(self.h, nsteps, check_every, ctypes.byref(pend)))
        return pend.value

    def poll(self, sid):
        st = L.SolveStatus()
        check(self._L.ausm_batch_poll(self.h, sid, ctypes.byref(st)))
        return {"state": STATES[st.state], "steps": st.steps, "res_l2": st.res_l2,
                "res_l2_0": st.res_l2_0, "cl": st.cl, "cd": st.cd}

    def state(self, sid):
        M, N = self._keep[sid]
        out = np.zeros((N + 3, M, 4), dtype=np.float32)
        check(self._L.ausm_batch_get_state(self.h, sid, ptr(out)))
        return out

    def release(self, sid):
        check(self._L.ausm_batch_release(self.h, sid))
        self._keep.pop(sid, None)
