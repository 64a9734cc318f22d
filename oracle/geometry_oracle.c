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
