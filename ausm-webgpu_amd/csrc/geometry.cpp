// Body and far-field generators of libausm (host code, FP64): the
// reference's host-side boundary generators (src/grid.js, src/airfoilDatParser.js)
// and the Kulfan CST and Bezier sections.  Each fills M points (x, y), counter-clockwise,
// in the point order the O-grid's body row (j = 0) or far-field row (j = N)
// expects.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ausm.h"

namespace ausm {
int set_error(int code, const char* fmt, ...);
}  // namespace ausm

using namespace ausm;

extern "C" {

int ausm_boundary_naca4(double m, double p, double t, double size, double x_offset,
                        double exponent, int spacing, int M, double* xy) {
  if (!xy || M < 4) return set_error(AUSM_ERR_ARG, "bad NACA arguments");
  if (m > 0 && (p <= 0 || p >= 1)) return set_error(AUSM_ERR_ARG, "camber position p must be in (0,1)");
  const double chord = 2.0 * size;
  for (int i = 0; i < M; ++i) {
    // generateNACA4Boundary (grid.js:38-52): parameter s = i / (M/2)
    const double s = (double)i / ((double)M / 2.0);
    const bool upper = s <= 1.0;
    double xc;
    if (spacing == AUSM_SPACING_COSINE)
      xc = upper ? 0.5 * (1.0 + std::cos(M_PI * s)) : 0.5 * (1.0 - std::cos(M_PI * (s - 1.0)));
    else
      xc = upper ? std::pow(1.0 - s, exponent) : std::pow(s - 1.0, exponent);
    const double yt = 5.0 * t *
                      (0.2969 * std::sqrt(xc) - 0.126 * xc - 0.3516 * xc * xc +
                       0.2843 * xc * xc * xc - 0.1036 * xc * xc * xc * xc);
    double yc = 0.0, dyc = 0.0;
    if (m > 0) {
      if (xc < p) {
        yc = (m / (p * p)) * (2.0 * p * xc - xc * xc);
        dyc = (m / (p * p)) * (2.0 * p - 2.0 * xc);
      } else {
        yc = (m / ((1.0 - p) * (1.0 - p))) * ((1.0 - 2.0 * p) + 2.0 * p * xc - xc * xc);
        dyc = (m / ((1.0 - p) * (1.0 - p))) * (2.0 * p - 2.0 * xc);
      }
    }
    // thickness applied perpendicular to the mean line
    const double th = std::atan(dyc);
    const double sx = yt * -std::sin(th), sy = yt / std::sqrt(1.0 + dyc * dyc);
    const double sgn = upper ? 1.0 : -1.0;
    xy[2 * i] = (xc - 0.5 + sgn * sx) * chord + x_offset;
    xy[2 * i + 1] = (yc + sgn * sy) * chord;
  }
  return AUSM_OK;
}

// Class-shape transformation section (Kulfan): with n = n_coef - 1,
//   y_u(x) = sqrt(x) (1 - x) sum_i A_u,i C(n,i) x^i (1-x)^(n-i) + x te_gap / 2
//   y_l(x) = sqrt(x) (1 - x) sum_i A_l,i C(n,i) x^i (1-x)^(n-i) - x te_gap / 2
// at the chordwise stations of the NACA generator (same point order).
static double cst_surface(const double* a, int n_coef, double x) {
  const int n = n_coef - 1;
  double s = 0.0, binom = 1.0;
  for (int i = 0; i <= n; ++i) {
    double xi = 1.0, xr = 1.0;
    for (int k = 0; k < i; ++k) xi *= x;
    for (int k = 0; k < n - i; ++k) xr *= 1.0 - x;
    s += a[i] * binom * xi * xr;
    binom = binom * (double)(n - i) / (double)(i + 1);
  }
  return std::sqrt(x) * (1.0 - x) * s;
}

int ausm_boundary_cst(const double* a_upper, const double* a_lower, int n_coef, double te_gap,
                      double size, double x_offset, double exponent, int spacing, int M,
                      double* xy) {
  if (!xy || !a_upper || !a_lower || M < 4) return set_error(AUSM_ERR_ARG, "bad CST arguments");
  if (n_coef < 1 || n_coef > 32) return set_error(AUSM_ERR_ARG, "CST needs 1..32 coefficients per side");
  const double chord = 2.0 * size;
  for (int i = 0; i < M; ++i) {
    const double s = (double)i / ((double)M / 2.0);
    const bool upper = s <= 1.0;
    double xc;
    if (spacing == AUSM_SPACING_COSINE)
      xc = upper ? 0.5 * (1.0 + std::cos(M_PI * s)) : 0.5 * (1.0 - std::cos(M_PI * (s - 1.0)));
    else
      xc = upper ? std::pow(1.0 - s, exponent) : std::pow(s - 1.0, exponent);
    const double y = upper ? cst_surface(a_upper, n_coef, xc) + xc * (0.5 * te_gap)
                           : cst_surface(a_lower, n_coef, xc) - xc * (0.5 * te_gap);
    xy[2 * i] = (xc - 0.5) * chord + x_offset;
    xy[2 * i + 1] = y * chord;
  }
  return AUSM_OK;
}

// Bezier section: upper and lower surfaces are Bezier curves of n_upper and
// n_lower control points (x, y) in chord units, each from the leading edge
// (t = 0) to the trailing edge (t = 1).  The curve parameter is spaced as the
// chordwise stations of the NACA generator (point order CCW from the upper
// trailing edge); points are evaluated by de Casteljau's construction.
static void bezier_point(const double* cp, int n, double t, double* x, double* y) {
  double bx[32], by[32];
  for (int k = 0; k < n; ++k) {
    bx[k] = cp[2 * k];
    by[k] = cp[2 * k + 1];
  }
  for (int r = n - 1; r > 0; --r)
    for (int k = 0; k < r; ++k) {
      bx[k] = (1.0 - t) * bx[k] + t * bx[k + 1];
      by[k] = (1.0 - t) * by[k] + t * by[k + 1];
    }
  *x = bx[0];
  *y = by[0];
}

int ausm_boundary_bezier(const double* upper, int n_upper, const double* lower, int n_lower,
                         double size, double x_offset, double exponent, int spacing, int M,
                         double* xy) {
  if (!xy || !upper || !lower || M < 4) return set_error(AUSM_ERR_ARG, "bad Bezier arguments");
  if (n_upper < 2 || n_upper > 32 || n_lower < 2 || n_lower > 32)
    return set_error(AUSM_ERR_ARG, "a Bezier surface needs 2..32 control points");
  for (int k = 0; k < 2 * n_upper; ++k)
    if (!std::isfinite(upper[k])) return set_error(AUSM_ERR_ARG, "non-finite Bezier control point");
  for (int k = 0; k < 2 * n_lower; ++k)
    if (!std::isfinite(lower[k])) return set_error(AUSM_ERR_ARG, "non-finite Bezier control point");
  if (upper[0] != lower[0] || upper[1] != lower[1])
    return set_error(AUSM_ERR_ARG, "the two surfaces must start at one leading-edge point");
  const double chord = 2.0 * size;
  for (int i = 0; i < M; ++i) {
    const double s = (double)i / ((double)M / 2.0);
    const bool up = s <= 1.0;
    double tc;
    if (spacing == AUSM_SPACING_COSINE)
      tc = up ? 0.5 * (1.0 + std::cos(M_PI * s)) : 0.5 * (1.0 - std::cos(M_PI * (s - 1.0)));
    else
      tc = up ? std::pow(1.0 - s, exponent) : std::pow(s - 1.0, exponent);
    double x, y;
    if (up) bezier_point(upper, n_upper, tc, &x, &y);
    else bezier_point(lower, n_lower, tc, &x, &y);
    xy[2 * i] = (x - 0.5) * chord + x_offset;
    xy[2 * i + 1] = y * chord;
  }
  return AUSM_OK;
}

int ausm_boundary_circle(double radius, int M, double* xy) {
  if (!xy || M < 4 || !(radius > 0)) return set_error(AUSM_ERR_ARG, "bad circle arguments");
  for (int i = 0; i < M; ++i) {
    const double t = (double)i / (double)M;  // grid.js:232-237
    xy[2 * i] = radius * std::cos(t * M_PI * 2.0);
    xy[2 * i + 1] = radius * std::sin(t * M_PI * 2.0);
  }
  return AUSM_OK;
}

int ausm_boundary_rect(double ratio, int M, double* xy) {
  if (!xy || M < 4 || !(ratio > 0)) return set_error(AUSM_ERR_ARG, "bad rectangle arguments");
  for (int i = 0; i < M; ++i) {
    double t = std::fmod((double)i / (double)M + 0.125, 1.0);  // grid.js:238-255
    double x, y;
    if (t < 0.25) {
      t *= 4.0;
      x = ratio;
      y = t * 2.0 - 1.0;
    } else if (t < 0.5) {
      t = (t - 0.25) * 4.0;
      x = ((1.0 - t) * 2.0 - 1.0) * ratio;
      y = 1.0;
    } else if (t < 0.75) {
      t = (t - 0.5) * 4.0;
      x = -ratio;
      y = (1.0 - t) * 2.0 - 1.0;
    } else {
      t = (t - 0.75) * 4.0;
      x = (t * 2.0 - 1.0) * ratio;
      y = -1.0;
    }
    xy[2 * i] = x;
    xy[2 * i + 1] = y;
  }
  return AUSM_OK;
}


// generateSearsHaackBoundary (src/grid.js:53-61): a Sears-Haack body of
// volume V and length L, r(s) = 8/pi sqrt(2V/(3L)) (s(1-s))^(3/4), nose at
// x_offset - L, base at x_offset; upper surface base -> nose, then lower.
int ausm_boundary_sears_haack(double V, double L, double x_offset, int M, double* xy) {
  if (!xy || M < 4 || !(V > 0) || !(L > 0)) return set_error(AUSM_ERR_ARG, "bad Sears-Haack arguments");
  const double k = 8.0 / M_PI * std::sqrt(2.0 * V / (3.0 * L));
  auto r = [&](double s) { return k * std::pow(s * (1.0 - s), 0.75); };
  for (int i = 0; i < M; ++i) {
    double t = (double)i / ((double)M / 2.0);
    if (t < 1.0) {
      xy[2 * i] = L * (-t) + x_offset;
      xy[2 * i + 1] = r(t);
    } else {
      t -= 1.0;
      xy[2 * i] = L * (t - 1.0) + x_offset;
      xy[2 * i + 1] = -r(t);
    }
  }
  return AUSM_OK;
}

// generateHaackProjectileBoundary (src/grid.js:75-118): a Haack-series
// nose cone (fineness F, radius R, parameter C) on a cylindrical body with
// a boat tail, points spread by arc-length fractions of the four sections
// (nose cone length approximated by the hypotenuse R sqrt(1 + 4F^2)).
int ausm_boundary_haack_projectile(double F, double R, double L, double Rb, double Lb, double C,
                                   double x_offset, int M, double* xy) {
  if (!xy || M < 4 || !(F > 0) || !(R > 0) || !(L > 0) || !(Lb > 0) || !(Lb < 1) || Rb < 0 || Rb > 1)
    return set_error(AUSM_ERR_ARG, "bad Haack projectile arguments");
  auto rn = [&](double T) {
    return R * std::sqrt((T - std::sin(2.0 * T) * 0.5 + C / 3.0 * std::pow(std::sin(T), 3.0)) / M_PI);
  };
  const double coneLength = R * std::hypot(1.0, 2.0 * F);
  const double FR = F * R;
  const double bodyLength = FR * L * (1.0 - Lb);
  const double btX = L * FR * Lb;
  const double btLength = std::hypot(btX, (1.0 - Rb) * R);
  const double btSlope = -(1.0 - Rb) * R / btX;
  const double tailLength = R * Rb;
  const double inv = 1.0 / (coneLength + bodyLength + btLength + tailLength);
  const double tailFrac = tailLength * inv;
  const double btFrac = btLength * inv + tailFrac;
  const double bodyFrac = bodyLength * inv + btFrac;
  for (int i = 0; i < M; ++i) {
    double t = (double)i / ((double)M / 2.0);
    double x, y;
    if (t < 1.0) {
      if (t < tailFrac) {
        t /= tailFrac;
        x = FR * L + x_offset;
        y = R * Rb * t;
      } else if (t < btFrac) {
        t = (t - tailFrac) / (btFrac - tailFrac);
        const double xb = FR * L - btX * t;
        x = xb + x_offset;
        y = R + btSlope * (xb - bodyLength);
      } else if (t < bodyFrac) {
        t = (t - btFrac) / (bodyFrac - btFrac);
        x = bodyLength * (1.0 - t) + x_offset;
        y = R;
      } else {
        t = (t - bodyFrac) / (1.0 - bodyFrac);
        x = 2.0 * F * R * (-t) + x_offset;
        y = rn(std::acos(2.0 * t - 1.0));
      }
    } else {
      t -= 1.0;
      if (t < 1.0 - bodyFrac) {
        t = t / (1.0 - bodyFrac);
        x = 2.0 * F * R * (t - 1.0) + x_offset;
        y = -rn(std::acos(1.0 - 2.0 * t));
      } else if (t < 1.0 - btFrac) {
        t = (t - (1.0 - bodyFrac)) / ((1.0 - btFrac) - (1.0 - bodyFrac));
        x = bodyLength * t + x_offset;
        y = -R;
      } else if (t < 1.0 - tailFrac) {
        t = (t - (1.0 - btFrac)) / ((1.0 - tailFrac) - (1.0 - btFrac));
        const double xb = FR * L - btX * (1.0 - t);
        x = xb + x_offset;
        y = -R - btSlope * (xb - bodyLength);
      } else {
        t = (t - (1.0 - tailFrac)) / tailFrac;
        x = FR * L + x_offset;
        y = -R * Rb * (1.0 - t);
      }
    }
    xy[2 * i] = x;
    xy[2 * i + 1] = y;
  }
  return AUSM_OK;
}

// generatePolygonBoundary (src/grid.js:224-231): a regular N-gon of
// inradius R (rotated by a), squashed by F in y; N = 400, F = 1 is the
// reference's circle body.
int ausm_boundary_polygon(double R, int N, double F, double a, double x_offset, int M, double* xy) {
  if (!xy || M < 4 || N < 3 || !(R > 0)) return set_error(AUSM_ERR_ARG, "bad polygon arguments");
  const double b = 2.0 * M_PI / N;
  for (int i = 0; i < M; ++i) {
    const double t = (double)i / (double)M * 2.0 * M_PI;
    const double r = R / std::cos(std::fmod(t + a + b, b) - M_PI / N);
    xy[2 * i] = r * std::cos(t) + x_offset;
    xy[2 * i + 1] = r * F * std::sin(t);
  }
  return AUSM_OK;
}

// ---- airfoil coordinate files (src/airfoilDatParser.js) ---------------------

// numeric lines grouped into blank-line separated blocks; header lines and
// "N. N." point-count lines (integers > 1 in both columns) are dropped
// (airfoilDatParser.js:29-57)
static std::vector<std::vector<std::pair<double, double>>> dat_blocks(const char* text) {
  std::vector<std::vector<std::pair<double, double>>> blocks;
  std::vector<std::pair<double, double>> cur;
  const char* p = text;
  while (*p) {
    const char* e = p;
    while (*e && *e != '\n') ++e;
    std::string line(p, e);
    p = *e ? e + 1 : e;
    size_t a = 0;
    while (a < line.size() && std::isspace((unsigned char)line[a])) ++a;
    if (a == line.size()) {
      if (!cur.empty()) blocks.push_back(cur), cur.clear();
      continue;
    }
    // the first two whitespace-separated tokens, parsed as numbers (a
    // token parses when it starts with a number, as JavaScript parseFloat)
    const char* s0 = line.c_str() + a;
    const char* t1 = s0;
    while (*t1 && !std::isspace((unsigned char)*t1)) ++t1;
    const char* s1 = t1;
    while (*s1 && std::isspace((unsigned char)*s1)) ++s1;
    if (!*s1) continue;
    char* end0 = nullptr;
    char* end1 = nullptr;
    const double x = std::strtod(s0, &end0);
    const double y = std::strtod(s1, &end1);
    if (end0 == s0 || end1 == s1 || !std::isfinite(x) || !std::isfinite(y)) continue;
    if (x == std::floor(x) && y == std::floor(y) && x > 1 && y > 1) continue;
    cur.push_back({x, y});
  }
  if (!cur.empty()) blocks.push_back(cur);
  return blocks;
}

// centripetal Catmull-Rom (alpha = 1/2), Barry-Goldman evaluation on the
// segment p1 -> p2 at local parameter u (airfoilDatParser.js:142-173)
static void catmull_rom(const double* p0, const double* p1, const double* p2, const double* p3,
                        double u, double* out) {
  auto knot = [](const double* a, const double* b) {
    const double dx = b[0] - a[0], dy = b[1] - a[1];
    return std::pow(dx * dx + dy * dy, 0.25);
  };
  const double t0 = 0.0;
  const double t1 = t0 + knot(p0, p1);
  const double t2 = t1 + knot(p1, p2);
  const double t3 = t2 + knot(p2, p3);
  const double t = t1 + u * (t2 - t1);
  auto lerp2 = [&](const double* a, const double* b, double ta, double tb, double* o) {
    if (std::fabs(tb - ta) < 1e-12) {
      o[0] = (a[0] + b[0]) / 2.0;
      o[1] = (a[1] + b[1]) / 2.0;
      return;
    }
    const double f = (t - ta) / (tb - ta);
    o[0] = a[0] + f * (b[0] - a[0]);
    o[1] = a[1] + f * (b[1] - a[1]);
  };
  double A1[2], A2[2], A3[2], B1[2], B2[2];
  lerp2(p0, p1, t0, t1, A1);
  lerp2(p1, p2, t1, t2, A2);
  lerp2(p2, p3, t2, t3, A3);
  lerp2(A1, A2, t0, t2, B1);
  lerp2(A2, A3, t1, t3, B2);
  lerp2(B1, B2, t1, t2, out);
}

// parseDat + parseAirfoil + sampleAirfoil (airfoilDatParser.js:67-293):
// Selig (one wrap-around list, split at the minimum x) or Lednicer (two
// blocks from the leading edge), joined upper TE -> LE -> lower TE, a
// Catmull-Rom spline over normalised arc length with reflected phantom end
// points, sampled at M logit-biased positions (denser at the nose) and
// mapped to chord 2*size at x_offset.
int ausm_boundary_dat(const char* text, double size, double x_offset, int M, double* xy,
                      int* format) {
  if (!text || !xy || M < 4) return set_error(AUSM_ERR_ARG, "bad airfoil-file arguments");
  const auto blocks = dat_blocks(text);
  if (blocks.empty()) return set_error(AUSM_ERR_ARG, "no numeric data in the airfoil file");
  std::vector<std::pair<double, double>> upper, lower;
  const bool lednicer = blocks.size() >= 2 && blocks[0][0].first < 0.1;
  if (lednicer) {
    upper.assign(blocks[0].rbegin(), blocks[0].rend());
    lower = blocks[1];
  } else {
    std::vector<std::pair<double, double>> pts;
    for (const auto& b : blocks) pts.insert(pts.end(), b.begin(), b.end());
    if (pts.size() < 4) return set_error(AUSM_ERR_ARG, "too few coordinate pairs in the airfoil file");
    size_t le = 0;
    for (size_t i = 1; i < pts.size(); ++i)
      if (pts[i].first < pts[le].first) le = i;
    upper.assign(pts.begin(), pts.begin() + le + 1);
    lower.assign(pts.begin() + le, pts.end());
  }
  if (upper.empty() || lower.empty()) return set_error(AUSM_ERR_ARG, "airfoil file has an empty surface");
  if (format) *format = lednicer ? 1 : 0;
  std::vector<std::pair<double, double>> c = upper;
  const bool dup = std::fabs(upper.back().first - lower.front().first) < 1e-9 &&
                   std::fabs(upper.back().second - lower.front().second) < 1e-9;
  c.insert(c.end(), lower.begin() + (dup ? 1 : 0), lower.end());
  const size_t n = c.size();
  if (n < 2) return set_error(AUSM_ERR_ARG, "too few coordinate pairs in the airfoil file");
  // control points with phantom ends, and normalised arc-length knots
  std::vector<double> cp(2 * (n + 2));
  cp[0] = 2.0 * c[0].first - c[1].first;
  cp[1] = 2.0 * c[0].second - c[1].second;
  for (size_t i = 0; i < n; ++i) {
    cp[2 * (i + 1)] = c[i].first;
    cp[2 * (i + 1) + 1] = c[i].second;
  }
  cp[2 * (n + 1)] = 2.0 * c[n - 1].first - c[n - 2].first;
  cp[2 * (n + 1) + 1] = 2.0 * c[n - 1].second - c[n - 2].second;
  std::vector<double> tn(n, 0.0);
  for (size_t i = 1; i < n; ++i) {
    const double dx = c[i].first - c[i - 1].first, dy = c[i].second - c[i - 1].second;
    tn[i] = tn[i - 1] + std::sqrt(dx * dx + dy * dy);
  }
  const double total = tn[n - 1];
  if (!(total > 0)) return set_error(AUSM_ERR_ARG, "degenerate airfoil contour");
  for (auto& v : tn) v = v / total;
  const double scale = 2.0 * size;
  const double a = 5.0;
  for (int k = 0; k < M; ++k) {
    const double xk = (double)k / (double)M;
    const double b = ((a - 1.0) * xk + 0.5) / a;
    double t = 0.5 + std::log(b / (1.0 - b)) / (2.0 * std::log(2.0 * a - 1.0));
    t = std::fmax(0.0, std::fmin(1.0, t));
    size_t lo = 0, hi = n - 2;
    while (lo < hi) {
      const size_t mid = (lo + hi) >> 1;
      if (tn[mid + 1] < t) lo = mid + 1;
      else hi = mid;
    }
    const double u = tn[lo + 1] > tn[lo] ? (t - tn[lo]) / (tn[lo + 1] - tn[lo]) : 0.0;
    double pt[2];
    catmull_rom(&cp[2 * lo], &cp[2 * (lo + 1)], &cp[2 * (lo + 2)], &cp[2 * (lo + 3)], u, pt);
    xy[2 * k] = (pt[0] - 0.5) * scale + x_offset;
    xy[2 * k + 1] = pt[1] * scale;
  }
  return AUSM_OK;
}

}  // extern "C"
