// EAM spline tables built on the host: the reference's interpolation of a
// tabulated function and the FP32 force tables of the mixed path.  Plain C++
// (no HIP), so the same code also builds into a CPU test of the tables.
#pragma once

#include <vector>

namespace cac {

// PairEAM::interpolate (pair_eam.cpp:586-614), f[1..n] -> spl[(n+1)*7]
inline void interpolate(int n, double delta, const std::vector<double>& f, std::vector<double>& spl) {
  spl.assign((size_t)(n + 1) * 7, 0.0);
  auto S = [&](int m, int k) -> double& { return spl[(size_t)m * 7 + k]; };
  for (int m = 1; m <= n; m++) S(m, 6) = f[m];
  S(1, 5) = S(2, 6) - S(1, 6);
  S(2, 5) = 0.5 * (S(3, 6) - S(1, 6));
  S(n - 1, 5) = 0.5 * (S(n, 6) - S(n - 2, 6));
  S(n, 5) = S(n, 6) - S(n - 1, 6);
  for (int m = 3; m <= n - 2; m++)
    S(m, 5) = ((S(m - 2, 6) - S(m + 2, 6)) + 8.0 * (S(m + 1, 6) - S(m - 1, 6))) / 12.0;
  for (int m = 1; m <= n - 1; m++) {
    S(m, 4) = 3.0 * (S(m + 1, 6) - S(m, 6)) - 2.0 * S(m, 5) - S(m + 1, 5);
    S(m, 3) = S(m, 5) + S(m + 1, 5) - 2.0 * (S(m + 1, 6) - S(m, 6));
  }
  S(n, 4) = 0.0;
  S(n, 3) = 0.0;
  for (int m = 1; m <= n; m++) {
    S(m, 2) = S(m, 5) / delta;
    S(m, 1) = 2.0 * S(m, 4) / delta;
    S(m, 0) = 3.0 * S(m, 3) / delta;
  }
}

// Force tables of the mixed path.  On each knot interval m of the reference's
// splines (r = (m - 1 + p) dr, 0 <= p <= 1) the FP32 pass evaluates three
// functions of r that the reference forms per pair in FP64 (pair_eam.cpp:
// 933-950, one element):
//   A = -rho'(r) / r, B = -phi'(r) / r with phi = z2 / r and
//   phi' = z2' / r - phi / r, so the pair force is (F'_i + F'_j) A + B;
//   C = phi(r), the pair energy of energy and virial steps,
// each as the cubic Hermite interpolant in p of the interval's own formula
// (values at p = 0 and 1, slopes by central differences 1e-4 wide).  Inside
// one interval the reference's splines are polynomials and 1 / r varies by
// dr / r, so the interpolant departs from the FP64 functions by about 1e-11
// of their range: the FP32 pass skips 1 / r, phi and phi' per pair and stays
// as close to the reference as the FP32 evaluation of its splines.  Interval
// nr holds the reference's clamp (knot nr - 1 at p = 1), interval nr + 1 is
// all zeros (an entry beyond the cutoff contributes an exact zero).  Planes:
// A, B, C, each nr + 2 float4 (a3, a2, a1, a0).
inline void mixed_force_table(int nr, double dr, const std::vector<double>& rs, const std::vector<double>& zs,
                              std::vector<float>& t) {
  auto fv = [&](int m, double p, double out[3]) {
    const double* cr = &rs[(size_t)m * 7];
    const double* cz = &zs[(size_t)m * 7];
    const double rhop = (cr[0] * p + cr[1]) * p + cr[2];
    const double z2 = ((cz[3] * p + cz[4]) * p + cz[5]) * p + cz[6];
    const double z2p = (cz[0] * p + cz[1]) * p + cz[2];
    const double r = (m - 1 + p) * dr;
    const double phi = z2 / r;
    out[0] = -rhop / r;
    out[1] = -(z2p / r - phi / r) / r;
    out[2] = phi;
  };
  const size_t plane = (size_t)(nr + 2) * 4;
  t.assign(3 * plane, 0.0f);
  const double h = 1e-4;
  for (int m = 1; m < nr; m++) {
    double v0[3], v1[3], a[3], b[3], c[3], d[3];
    fv(m, 0.0, v0);
    fv(m, 1.0, v1);
    fv(m, h, a);
    fv(m, -h, b);
    fv(m, 1.0 + h, c);
    fv(m, 1.0 - h, d);
    for (int w = 0; w < 3; w++) {
      const double g0 = (a[w] - b[w]) / (2.0 * h), g1 = (c[w] - d[w]) / (2.0 * h);
      float* o = &t[w * plane + (size_t)m * 4];
      o[0] = (float)(2.0 * (v0[w] - v1[w]) + g0 + g1);
      o[1] = (float)(3.0 * (v1[w] - v0[w]) - 2.0 * g0 - g1);
      o[2] = (float)g0;
      o[3] = (float)v0[w];
    }
  }
  double vend[3];
  fv(nr - 1, 1.0, vend);
  for (int w = 0; w < 3; w++) t[w * plane + (size_t)nr * 4 + 3] = (float)vend[w];
}

}  // namespace cac
