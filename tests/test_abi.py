"""CPU checks of the C ABI: libausm.so loads, exports every function that
include/ausm.h declares, and its host-only geometry generators agree with the
oracle's restatement of the reference generators (src/grid.js)."""
import os
import re

import numpy as np
import pytest

from oracle_lib import circle_boundary as o_circle
from oracle_lib import naca4_boundary as o_naca
from oracle_lib import rect_boundary as o_rect

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ausm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ausm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for n in ("ausm_ctx_create", "ausm_step", "ausm_get_state", "ausm_forces",
              "ausm_batch_submit", "ausm_batch_advance"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import ausm
    from ausm import lib as L

    lib = ausm.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in L.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.ausm_version() == 1


def test_errors_without_gpu_are_reported_not_raised():
    import ctypes

    import ausm

    lib = ausm.load()
    h = ctypes.c_void_p()
    rc = lib.ausm_ctx_create(0, 2, 2, ctypes.byref(h))
    assert rc == -1 and b"out of range" in lib.ausm_last_error()


@pytest.mark.parametrize("code,spacing", [(12, "power"), (2412, "power"), (4412, "cosine")])
def test_naca_generator_matches_oracle(code, spacing):
    from ausm import naca4_boundary, naca4_digits

    M = 256
    a = naca4_boundary(M, code=code, size=0.3, x_offset=-0.7, exponent=1.4, spacing=spacing)
    m, p, t = naca4_digits(code)
    b = o_naca(M, m, p, t, size=0.3, x_offset=-0.7, expo=1.4, spacing=1 if spacing == "cosine" else 0)
    assert np.allclose(a, b, rtol=0, atol=1e-15)


@pytest.mark.parametrize("spacing", ["power", "cosine"])
def test_bezier_generator_matches_oracle(spacing):
    """The library's de Casteljau evaluation against the oracle's Bernstein
    form, and its argument checks."""
    import ausm
    from ausm import bezier_boundary

    from oracle_lib import bezier_boundary as o_bez

    cu = [[0.0, 0.0], [0.0, 0.06], [0.25, 0.1], [0.6, 0.08], [0.85, 0.03], [1.0, 0.0]]
    cl = [[0.0, 0.0], [0.0, -0.05], [0.3, -0.06], [0.7, -0.01], [1.0, 0.0]]
    a = bezier_boundary(512, cu, cl, size=0.5, x_offset=0.1, spacing=spacing)
    b = o_bez(512, cu, cl, size=0.5, x_offset=0.1, spacing=1 if spacing == "cosine" else 0)
    assert np.allclose(a, b, rtol=0, atol=1e-15)
    with pytest.raises(ausm.AusmError, match="leading-edge"):
        bezier_boundary(64, cu, [[0.0, 0.001], [1.0, 0.0]])
    with pytest.raises(ausm.AusmError, match="2..32"):
        bezier_boundary(64, cu, [[0.0, 0.0]])
    with pytest.raises(ausm.AusmError, match="non-finite"):
        bezier_boundary(64, cu, [[0.0, 0.0], [float("nan"), 0.0]])


def test_far_field_generators_match_oracle():
    from ausm import circle_boundary, rect_boundary

    assert np.allclose(circle_boundary(100, 20.0), o_circle(100, 20.0), atol=1e-14)
    assert np.allclose(rect_boundary(512, 1.5), o_rect(512, 1.5), atol=1e-15)


@pytest.mark.parametrize("spacing", ["power", "cosine"])
def test_cst_generator_matches_oracle(spacing):
    from ausm import cst_boundary

    from oracle_lib import cst_boundary as o_cst

    au = [0.13, 0.13, 0.16, 0.14, 0.18, 0.20]
    al = [-0.13, -0.14, -0.10, -0.03, 0.03, 0.06]
    a = cst_boundary(512, au, al, te_gap=0.002, size=0.5, x_offset=0.1, spacing=spacing)
    b = o_cst(512, au, al, te_gap=0.002, size=0.5, x_offset=0.1,
              spacing=1 if spacing == "cosine" else 0)
    assert np.allclose(a, b, rtol=0, atol=1e-15)
