"""CPU ORACLE — test infrastructure only.

Pure-Python restatement of the reference's airfoil coordinate-file front-end
(huj31415/ausm-webgpu src/airfoilDatParser.js), the checker for the
library's `ausm_boundary_dat`.  Only tests import it.

  blocks(text)              extractBlocks, airfoilDatParser.js:29-57
  surfaces(text)            parseDat, airfoilDatParser.js:67-109
  contour(text)             parseAirfoil's join, airfoilDatParser.js:247-265
  Spline(points)(t)         buildSpline + catmullRom, airfoilDatParser.js:142-233
  sample_positions(n)       sampleAirfoil's logit bias, airfoilDatParser.js:276-292
  sample(text, n, size, x0) sampleAirfoil, airfoilDatParser.js:276-293

Pinned in tests/test_geometry.py: the spline passes through its control
points, reproduces straight lines, the two file formats of one section give
one contour, the sample positions are symmetric about 1/2, a symmetric
section samples symmetrically, and SC(2)-0012's sampled thickness is 12 %.
"""
import math


def _number(tok):
    """Leading numeric prefix of a token, as JavaScript parseFloat reads it
    (None when there is none)."""
    best = None
    for k in range(1, len(tok) + 1):
        try:
            v = float(tok[:k])
        except ValueError:
            continue
        if tok[:k].strip().lower() in ("inf", "nan", "infinity", "+inf", "-inf"):
            continue
        best = v
    return best


def blocks(text):
    out, cur = [], []
    for line in text.split("\n"):
        tr = line.strip()
        if not tr:
            if cur:
                out.append(cur)
                cur = []
            continue
        toks = tr.split()
        if len(toks) < 2:
            continue
        x, y = _number(toks[0]), _number(toks[1])
        if x is None or y is None:
            continue
        if float(x).is_integer() and float(y).is_integer() and x > 1 and y > 1:
            continue
        cur.append((x, y))
    if cur:
        out.append(cur)
    return out


def surfaces(text):
    """(upper TE->LE, lower LE->TE, format)"""
    b = blocks(text)
    if not b:
        raise ValueError("no numeric data")
    if len(b) >= 2 and b[0][0][0] < 0.1:
        return list(reversed(b[0])), list(b[1]), "lednicer"
    pts = [p for blk in b for p in blk]
    if len(pts) < 4:
        raise ValueError("too few points")
    le = 0
    for i in range(1, len(pts)):
        if pts[i][0] < pts[le][0]:
            le = i
    return pts[:le + 1], pts[le:], "selig"


def contour(text):
    up, lo, _ = surfaces(text)
    dup = abs(up[-1][0] - lo[0][0]) < 1e-9 and abs(up[-1][1] - lo[0][1]) < 1e-9
    return up + (lo[1:] if dup else lo)


def _catmull_rom(p0, p1, p2, p3, u):
    def knot(a, b):
        dx, dy = b[0] - a[0], b[1] - a[1]
        return (dx * dx + dy * dy) ** 0.25

    t0 = 0.0
    t1 = t0 + knot(p0, p1)
    t2 = t1 + knot(p1, p2)
    t3 = t2 + knot(p2, p3)
    t = t1 + u * (t2 - t1)

    def lerp(a, b, ta, tb):
        if abs(tb - ta) < 1e-12:
            return ((a[0] + b[0]) / 2, (a[1] + b[1]) / 2)
        f = (t - ta) / (tb - ta)
        return (a[0] + f * (b[0] - a[0]), a[1] + f * (b[1] - a[1]))

    a1 = lerp(p0, p1, t0, t1)
    a2 = lerp(p1, p2, t1, t2)
    a3 = lerp(p2, p3, t2, t3)
    b1 = lerp(a1, a2, t0, t2)
    b2 = lerp(a2, a3, t1, t3)
    return lerp(b1, b2, t1, t2)


class Spline:
    """Centripetal Catmull-Rom through `pts` over normalised arc length, with
    phantom end points reflected through the first and last points."""

    def __init__(self, pts):
        n = len(pts)
        self.cp = ([(2 * pts[0][0] - pts[1][0], 2 * pts[0][1] - pts[1][1])] + list(pts) +
                   [(2 * pts[-1][0] - pts[-2][0], 2 * pts[-1][1] - pts[-2][1])])
        s = [0.0]
        for i in range(1, n):
            s.append(s[-1] + math.hypot(pts[i][0] - pts[i - 1][0], pts[i][1] - pts[i - 1][1]))
        self.t = [v / s[-1] for v in s]

    def __call__(self, t):
        t = max(0.0, min(1.0, t))
        tn = self.t
        seg = 0
        while seg < len(tn) - 2 and tn[seg + 1] < t:
            seg += 1
        u = (t - tn[seg]) / (tn[seg + 1] - tn[seg]) if tn[seg + 1] > tn[seg] else 0.0
        c = self.cp
        return _catmull_rom(c[seg], c[seg + 1], c[seg + 2], c[seg + 3], u)


def sample_positions(n, a=5.0):
    out = []
    for k in range(n):
        x = k / n
        b = ((a - 1) * x + 0.5) / a
        out.append(0.5 + math.log(b / (1 - b)) / (2 * math.log(2 * a - 1)))
    return out


def sample(text, n, size=0.5, x_offset=0.0):
    sp = Spline(contour(text))
    scale = 2.0 * size
    pts = []
    for t in sample_positions(n):
        x, y = sp(t)
        pts.append(((x - 0.5) * scale + x_offset, y * scale))
    return pts
