"""Body generators and the airfoil coordinate-file front-end (host code of
libausm, no GPU): equal to the oracle's restatements and pinned by what the
geometry fixes (volumes, support functions, symmetry, interpolation)."""
import math
import os

import numpy as np
import pytest

import ausm
from oracle_lib import ROOT, haack_boundary, polygon_boundary, sears_haack_boundary

import sys

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import airfoil_dat as dat_oracle  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def revolve_volume(xy):
    """Volume of revolution about the x axis of the upper half (y >= 0) of a
    closed contour, by the trapezoid rule in x."""
    up = xy[xy[:, 1] >= 0]
    up = up[np.argsort(up[:, 0])]
    x, r = up[:, 0], up[:, 1]
    return float(np.sum(np.pi * 0.5 * (r[1:] ** 2 + r[:-1] ** 2) * np.diff(x)))


def test_sears_haack_matches_oracle_and_volume():
    a = ausm.sears_haack_boundary(512, V=0.0002, L=0.4, x_offset=-0.7)
    assert np.allclose(a, sears_haack_boundary(512, 0.0002, 0.4, -0.7), rtol=0, atol=1e-15)
    # the Sears-Haack radius law encloses exactly the volume V
    b = ausm.sears_haack_boundary(8192, V=0.0002, L=0.4, x_offset=0.0)
    assert math.isclose(revolve_volume(b), 0.0002, rel_tol=2e-3)
    # upper and lower surfaces mirror each other
    M = 512
    assert np.allclose(a[1:M // 2, 0], a[M - 1:M // 2:-1, 0], atol=1e-15)
    assert np.allclose(a[1:M // 2, 1], -a[M - 1:M // 2:-1, 1], atol=1e-15)


@pytest.mark.parametrize("C", [0.0, 1.0 / 3.0])
def test_haack_projectile_matches_oracle_and_nose(C):
    M = 1024
    a = ausm.haack_projectile_boundary(M, C=C)
    assert np.allclose(a, haack_boundary(M, C=C), rtol=0, atol=1e-14)
    # mirror symmetry and a closed, base-to-nose-to-base contour
    assert np.allclose(a[1:M // 2, 0], a[M - 1:M // 2:-1, 0], atol=1e-14)
    assert np.allclose(a[1:M // 2, 1], -a[M - 1:M // 2:-1, 1], atol=1e-14)
    F, R = 6.0, 0.03
    assert a[:, 1].max() <= R + 1e-15
    assert math.isclose(a[:, 0].min(), -0.7 - 2 * F * R, abs_tol=2 * F * R * 1e-3)
    if C == 0.0:
        # Von Karman (LD-Haack) nose: volume = half the circumscribing cylinder
        b = ausm.haack_projectile_boundary(16384, C=0.0, x_offset=0.0)
        nose = b[b[:, 0] <= 0]
        vol = revolve_volume(nose)
        assert math.isclose(vol, 0.5 * math.pi * R * R * (2 * F * R), rel_tol=3e-3)


@pytest.mark.parametrize("N,a", [(4, 0.0), (6, 0.3), (400, 0.0)])
def test_polygon_matches_oracle_and_support_function(N, a):
    M, R = 600, 0.1
    p = ausm.polygon_boundary(M, R=R, N=N, F=1.0, a=a, x_offset=-0.7)
    assert np.allclose(p, polygon_boundary(M, R, N, 1.0, a, -0.7), rtol=0, atol=1e-15)
    # every point lies on the N-gon of inradius R: its largest projection on
    # the edge normals is R
    q = p - np.array([-0.7, 0.0])
    ang = np.pi / N - a + 2 * np.pi * np.arange(N) / N
    proj = q @ np.stack([np.cos(ang), np.sin(ang)])
    assert np.allclose(proj.max(axis=1), R, rtol=0, atol=1e-13)
    # F squashes y only
    s = ausm.polygon_boundary(M, R=R, N=N, F=0.1, a=a, x_offset=-0.7)
    assert np.allclose(s[:, 0], p[:, 0]) and np.allclose(s[:, 1], 0.1 * p[:, 1])


def test_dat_oracle_spline_pins():
    # the spline passes through its control points and keeps straight lines
    pts = [(1.0, 0.0), (0.7, 0.05), (0.3, 0.06), (0.1, 0.04), (0.0, 0.0), (0.1, -0.03),
           (0.5, -0.04), (1.0, 0.0)]
    sp = dat_oracle.Spline(pts)
    for t, p in zip(sp.t, pts):
        x, y = sp(t)
        assert abs(x - p[0]) < 1e-12 and abs(y - p[1]) < 1e-12
    line = [(1.0 - 0.1 * k, 0.5 * (1.0 - 0.1 * k) + 0.2) for k in range(11)]
    sl = dat_oracle.Spline(line)
    for t in np.linspace(0, 1, 37):
        x, y = sl(t)
        assert abs(y - (0.5 * x + 0.2)) < 1e-12
    # logit-biased positions: from the upper TE (0) round to the lower TE,
    # symmetric about 1/2
    s = dat_oracle.sample_positions(512)
    assert abs(s[0]) < 1e-15 and abs(s[256] - 0.5) < 1e-15
    assert all(abs(s[k] + s[512 - k] - 1.0) < 1e-14 for k in range(1, 512))


@pytest.mark.parametrize("name,fmt", [("sc20012.dat", "lednicer"), ("bacxxx.dat", "lednicer"),
                                      ("kc135winglet.dat", "selig")])
def test_dat_front_end_matches_oracle(name, fmt):
    text = read(name)
    a, f = ausm.dat_boundary(512, text, size=0.3, x_offset=-0.7)
    assert f == fmt
    b = np.array(dat_oracle.sample(text, 512, 0.3, -0.7))
    assert np.allclose(a, b, rtol=0, atol=1e-14)


def test_dat_formats_give_one_contour_and_sc20012_is_12_percent():
    text = read("sc20012.dat")
    up, lo, fmt = dat_oracle.surfaces(text)
    assert fmt == "lednicer"
    # the same section written as a Selig file
    selig = "SC(2)-0012 as Selig\n" + "\n".join(f"{x:.6f} {y:.6f}" for x, y in up + lo[1:])
    a, fa = ausm.dat_boundary(400, text, size=0.5, x_offset=0.0)
    b, fb = ausm.dat_boundary(400, selig, size=0.5, x_offset=0.0)
    assert (fa, fb) == ("lednicer", "selig")
    assert np.allclose(a, b, rtol=0, atol=1e-12)
    # symmetric section: point k and point M - k mirror each other; 12 % thick
    M = 400
    assert np.allclose(a[1:M // 2, 0], a[M - 1:M // 2:-1, 0], atol=1e-9)
    assert np.allclose(a[1:M // 2, 1], -a[M - 1:M // 2:-1, 1], atol=1e-9)
    assert abs(2 * a[:, 1].max() - 0.12) < 5e-4
    # the first point is the upper trailing edge of the file
    assert abs(a[0, 0] - (up[0][0] - 0.5)) < 1e-12 and abs(a[0, 1] - up[0][1]) < 1e-12


def test_dat_rejects_empty_input():
    with pytest.raises(ausm.AusmError):
        ausm.dat_boundary(64, "just a title\n\n")
