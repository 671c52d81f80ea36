// Real spherical harmonics shared by the host-side builders.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace fdmgpu {

// Real spherical harmonics Y_lm(r^) for l <= lmax, index l*l + l + m,
// tesseral convention without the Condon-Shortley phase (the reference's
// cylm + ylmcr, general.f90:11606-11720).  Normalised associated Legendre
// functions by the standard three-term recurrence in l.
inline void real_ylm(int lmax, double x, double y, double z, double *out) {
  const double r = std::sqrt(x * x + y * y + z * z);
  const double rho = std::sqrt(x * x + y * y);
  const double ct = (r > 1e-6) ? z / r : 1.0;
  double st = 0;
  std::complex<double> eph(1.0, 0.0);
  if (rho > 1e-6) {
    st = std::sqrt(1 - ct * ct);
    eph = std::complex<double>(x, y) / rho;
  }
  const int nl = lmax + 1;
  std::vector<double> P((size_t)nl * nl, 0.0);  // P[l*nl + m]
  P[0] = 1.0 / std::sqrt(4 * 3.1415926535897932384626433832);
  for (int m = 1; m <= lmax; m++) P[m * nl + m] = std::sqrt((2 * m + 1.0) / (2.0 * m)) * st * P[(m - 1) * nl + m - 1];
  for (int m = 0; m < lmax; m++) P[(m + 1) * nl + m] = std::sqrt(2 * m + 3.0) * ct * P[m * nl + m];
  for (int m = 0; m <= lmax; m++)
    for (int l = m + 2; l <= lmax; l++) {
      const double a = std::sqrt((4.0 * l * l - 1) / ((double)l * l - (double)m * m));
      const double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1));
      P[l * nl + m] = a * (ct * P[(l - 1) * nl + m] - b * P[(l - 2) * nl + m]);
    }
  std::complex<double> em(1.0, 0.0);
  std::vector<std::complex<double>> pw(nl);
  for (int m = 0; m <= lmax; m++) {
    pw[m] = em;
    em *= eph;
  }
  for (int l = 0; l <= lmax; l++) {
    out[l * l + l] = P[l * nl];
    for (int m = 1; m <= l; m++) {
      const double v = std::sqrt(2.0) * P[l * nl + m];
      out[l * l + l + m] = (st > 0) ? v * pw[m].real() : 0.0;
      out[l * l + l - m] = (st > 0) ? v * pw[m].imag() : 0.0;
    }
  }
}


}  // namespace fdmgpu
