"""ausm: MI355X-native explicit O-grid solver step (SLAU2 / SLAU / AUSM+-up,
MUSCL, TVD-RK3) behind the C ABI of include/ausm.h."""
from .lib import (AUSMUP, DT_GLOBAL, DT_LOCAL, GRID_JACOBI, GRID_SOR, INIT_FREESTREAM, INIT_REST,
                  MATH_EXACT, MATH_FAST, MINMOD, NOLIMITER, PREC_FP32, PREC_FP32C, PREC_FP64, SLAU, SLAU2,
                  VANALBADA, VANALBADA_CLAMPED, VANLEER, AusmError, load)
from .batch import Batch
from .solver import (Solver, bezier_boundary, circle_boundary, cst_boundary, dat_boundary, freestream,
                     haack_projectile_boundary, naca4_boundary, naca4_digits, polygon_boundary,
                     rect_boundary, sears_haack_boundary)

__all__ = ["Solver", "Batch", "AusmError", "load", "naca4_boundary", "naca4_digits", "cst_boundary", "bezier_boundary",
           "sears_haack_boundary",
           "haack_projectile_boundary", "polygon_boundary", "dat_boundary", "circle_boundary",
           "rect_boundary", "freestream", "SLAU2", "SLAU", "AUSMUP", "VANALBADA", "MINMOD",
           "VANLEER", "NOLIMITER", "VANALBADA_CLAMPED", "MATH_EXACT", "MATH_FAST", "INIT_REST", "INIT_FREESTREAM",
           "GRID_JACOBI", "GRID_SOR", "PREC_FP32", "PREC_FP64", "PREC_FP32C", "DT_GLOBAL", "DT_LOCAL"]
