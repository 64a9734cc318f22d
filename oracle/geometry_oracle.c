/*
 * CPU ORACLE — test infrastructure only.
 *
 * Double-precision restatement of the reference's host-side boundary
 * generators (src/grid.js), used by tests to check the library's own
 * generators.  JavaScript numbers are IEEE doubles, so the arithmetic is
 * done in double here too.
 */
#include <math.h>

/* generateNACA4Boundary (src/grid.js:38-52) for vertex index i of M,
 * generalised to continuous (m, p, t) and to cosine spacing:
 *   spacing 0: x = (1-s)^expo on the upper side, (s-1)^expo on the lower
 *              side, s = i/(M/2) (the reference's power law);
 *   spacing 1: x = (1+cos(pi*s))/2 upper, (1-cos(pi*(s-1)))/2 lower.
 * Chord = 2*size, mid-chord at x_offset; CCW from the trailing edge. */
void geo_naca4_point(double m, double p, double thick, double size, double x_offset,
                     double expo, int spacing, int M, int i, double* xy) {
  double t = (double)i / ((double)M / 2.0);
  size *= 2.0;
  const int upper = t <= 1.0;
  double xc;
  if (spacing == 0) xc = upper ? pow(1.0 - t, expo) : pow(t - 1.0, expo);
  else xc = upper ? 0.5 * (1.0 + cos(M_PI * t)) : 0.5 * (1.0 - cos(M_PI * (t - 1.0)));
  const double yt = 5.0 * thick *
                    (0.2969 * sqrt(xc) - 0.126 * xc - 0.3516 * xc * xc + 0.2843 * xc * xc * xc -
                     0.1036 * xc * xc * xc * xc);
  double yc, dyc;
  if (xc < p) {
    yc = (m / (p * p)) * (2.0 * p * xc - xc * xc);
    dyc = (m / (p * p)) * (2.0 * p - 2.0 * xc);
  } else {
    yc = (m / ((1.0 - p) * (1.0 - p))) * ((1.0 - 2.0 * p) + 2.0 * p * xc - xc * xc);
    dyc = (m / ((1.0 - p) * (1.0 - p))) * (2.0 * p - 2.0 * xc);
  }
  const double th = atan(dyc);
  const double sx = yt * -sin(th), sy = yt / sqrt(1.0 + dyc * dyc);
  if (upper) {
    xy[0] = (xc - 0.5 + sx) * size + x_offset;
    xy[1] = (yc + sy) * size;
  } else {
    xy[0] = (xc - 0.5 - sx) * size + x_offset;
    xy[1] = (yc - sy) * size;
  }
}

/* Kulfan class-shape transformation section (no reference counterpart;
 * the reference draws NACA 4-digit bodies only).  Shape function evaluated
 * from its definition: Bernstein polynomials of degree n = ncoef-1 with
 * binomial coefficients from a factorial table, class function
 * x^0.5 (1-x)^1; chord stations and placement as geo_naca4_point.  Pinned
 * by the mathematics (partition of unity, degree elevation, nose radius,
 * symmetry, trailing-edge gap) in tests/test_oracle_pins.py. */
static double geo_cst_y(const double* a, int ncoef, double x) {
  const int n = ncoef - 1;
  double fact[33];
  fact[0] = 1.0;
  for (int k = 1; k <= n; ++k) fact[k] = fact[k - 1] * (double)k;
  double s = 0.0;
  for (int i = 0; i <= n; ++i)
    s += a[i] * (fact[n] / (fact[i] * fact[n - i])) * pow(x, (double)i) * pow(1.0 - x, (double)(n - i));
  return pow(x, 0.5) * pow(1.0 - x, 1.0) * s;
}

void geo_cst_point(const double* au, const double* al, int ncoef, double te_gap, double size,
                   double x_offset, double expo, int spacing, int M, int i, double* xy) {
  const double t = (double)i / ((double)M / 2.0);
  const double chord = 2.0 * size;
  const int upper = t <= 1.0;
  double xc;
  if (spacing == 0) xc = upper ? pow(1.0 - t, expo) : pow(t - 1.0, expo);
  else xc = upper ? 0.5 * (1.0 + cos(M_PI * t)) : 0.5 * (1.0 - cos(M_PI * (t - 1.0)));
  const double y = upper ? geo_cst_y(au, ncoef, xc) + 0.5 * te_gap * xc
                         : geo_cst_y(al, ncoef, xc) - 0.5 * te_gap * xc;
  xy[0] = (xc - 0.5) * chord + x_offset;
  xy[1] = y * chord;
}

/* Bezier section (no counterpart in the reference): each surface is
 * evaluated in Bernstein form, sum_k C(n-1,k) t^k (1-t)^(n-1-k) P_k, with the
 * curve parameter t at the chordwise stations of geo_naca4_point; the upper
 * control polygon serves s <= 1, the lower one s > 1.  Pinned by the
 * mathematics (endpoint interpolation, linear precision, the parabola of a
 * quadratic curve, degree elevation, mirror symmetry) in
 * tests/test_oracle_pins.py. */
static void geo_bernstein(const double* cp, int n, double t, double* x, double* y) {
  const int d = n - 1;
  double fact[33];
  fact[0] = 1.0;
  for (int k = 1; k <= d; ++k) fact[k] = fact[k - 1] * (double)k;
  double sx = 0.0, sy = 0.0;
  for (int k = 0; k <= d; ++k) {
    const double b = (fact[d] / (fact[k] * fact[d - k])) * pow(t, (double)k) * pow(1.0 - t, (double)(d - k));
    sx += b * cp[2 * k];
    sy += b * cp[2 * k + 1];
  }
  *x = sx;
  *y = sy;
}

void geo_bezier_point(const double* cu, int nu, const double* cl, int nl, double size,
                      double x_offset, double expo, int spacing, int M, int i, double* xy) {
  const double s = (double)i / ((double)M / 2.0);
  const double chord = 2.0 * size;
  const int upper = s <= 1.0;
  double tc;
  if (spacing == 0) tc = upper ? pow(1.0 - s, expo) : pow(s - 1.0, expo);
  else tc = upper ? 0.5 * (1.0 + cos(M_PI * s)) : 0.5 * (1.0 - cos(M_PI * (s - 1.0)));
  double x, y;
  if (upper) geo_bernstein(cu, nu, tc, &x, &y);
  else geo_bernstein(cl, nl, tc, &x, &y);
  xy[0] = (x - 0.5) * chord + x_offset;
  xy[1] = y * chord;
}

/* generateCircularOuterBoundary (src/grid.js:232-237) */
void geo_circle_point(double radius, int M, int i, double* xy) {
  const double t = (double)i / (double)M;
  xy[0] = radius * cos(t * M_PI * 2.0);
  xy[1] = radius * sin(t * M_PI * 2.0);
}

/* generateRectangularOuterBoundary (src/grid.js:238-255) */
void geo_rect_point(double ratio, int M, int i, double* xy) {
  double t = fmod((double)i / (double)M + 0.125, 1.0);
  if (t < 0.25) {
    t *= 4.0;
    xy[0] = ratio;
    xy[1] = t * 2.0 - 1.0;
  } else if (t < 0.5) {
    t = (t - 0.25) * 4.0;
    xy[0] = ((1.0 - t) * 2.0 - 1.0) * ratio;
    xy[1] = 1.0;
  } else if (t < 0.75) {
    t = (t - 0.5) * 4.0;
    xy[0] = -ratio;
    xy[1] = (1.0 - t) * 2.0 - 1.0;
  } else {
    t = (t - 0.75) * 4.0;
    xy[0] = (t * 2.0 - 1.0) * ratio;
    xy[1] = -1.0;
  }
}

/* generateSearsHaackBoundary (src/grid.js:53-61) for vertex i of M. */
void geo_sears_haack_point(double V, double L, double x_offset, int M, int i, double* xy) {
  double t = (double)i / ((double)M / 2.0);
  const double c = 8.0 / M_PI * sqrt(2.0 * V / (3.0 * L));
  if (t < 1.0) {
    xy[0] = -L * t + x_offset;
    xy[1] = c * pow(t * (1.0 - t), 0.75);
  } else {
    t = t - 1.0;
    xy[0] = L * (t - 1.0) + x_offset;
    xy[1] = -c * pow(t * (1.0 - t), 0.75);
  }
}

/* generateHaackProjectileBoundary (src/grid.js:75-118) for vertex i of M:
 * tail cap, boat tail, cylinder and Haack nose on the upper side (TE to
 * nose), mirrored on the lower side. */
static double haack_r(double R, double C, double T) {
  const double s = sin(T);
  return R * sqrt((T - 0.5 * sin(2.0 * T) + C / 3.0 * s * s * s) / M_PI);
}

void geo_haack_point(double F, double R, double L, double Rb, double Lb, double C, double x_offset,
                     int M, int i, double* xy) {
  const double FR = F * R;
  const double cone = R * sqrt(1.0 + 4.0 * F * F);
  const double body = FR * L * (1.0 - Lb);
  const double btx = FR * L * Lb;
  const double bt = sqrt(btx * btx + (1.0 - Rb) * (1.0 - Rb) * R * R);
  const double slope = -(1.0 - Rb) * R / btx;
  const double tail = R * Rb;
  const double tot = cone + body + bt + tail;
  const double f1 = tail / tot, f2 = bt / tot + f1, f3 = body / tot + f2;
  double t = (double)i / ((double)M / 2.0);
  const int up = t < 1.0;
  /* position along the upper traversal (TE -> nose) */
  double u = up ? t : 1.0 - (t - 1.0);
  double x, y;
  if (up ? (u < f1) : (u <= f1)) {
    const double s = u / f1;
    x = FR * L; y = R * Rb * s;
  } else if (up ? (u < f2) : (u <= f2)) {
    const double s = (u - f1) / (f2 - f1);
    x = FR * L - btx * s; y = R + slope * (x - body);
  } else if (up ? (u < f3) : (u <= f3)) {
    const double s = (u - f2) / (f3 - f2);
    x = body * (1.0 - s); y = R;
  } else {
    const double s = (u - f3) / (1.0 - f3);
    x = -2.0 * FR * s; y = haack_r(R, C, acos(2.0 * s - 1.0));
  }
  xy[0] = x + x_offset;
  xy[1] = up ? y : -y;
}

/* generatePolygonBoundary (src/grid.js:224-231) for vertex i of M. */
void geo_polygon_point(double R, int N, double F, double a, double x_offset, int M, int i,
                       double* xy) {
  const double t = 2.0 * M_PI * (double)i / (double)M;
  const double b = 2.0 * M_PI / (double)N;
  const double ph = fmod(t + a + b, b) - M_PI / (double)N;
  const double r = R / cos(ph);
  xy[0] = r * cos(t) + x_offset;
  xy[1] = F * r * sin(t);
}
