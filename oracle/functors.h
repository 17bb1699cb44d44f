// E_loc functors of the Hamiltonians with complex wave-function ratios --
// test infrastructure (see oracle.h).  TFIsing.hpp:28-39, XXZ.hpp:37-49 and
// J1J2.hpp:35-55: the diagonal term and the prefactor of every bond's ratio,
// accumulated in the functors' order.  ratio1(k) = psi(sigma^k)/psi(sigma),
// ratio2(k, l) = psi(sigma^{k,l})/psi(sigma).
#pragma once
#include <complex>

#include "oracle.h"

namespace oracle_detail {

template <class R1, class R2>
std::complex<double> eloc_functor(const or_ham* H, const int8_t* s, R1&& ratio1, R2&& ratio2) {
  const int n = H->n;
  std::complex<double> e = 0.0;
  if (H->kind == OR_TFI) {
    for (int i = 0; i < n; ++i) {
      e += H->J * (double)s[i] * (double)s[(i + 1) % n];
      e += H->h * ratio1(i);
    }
  } else if (H->kind == OR_XXZ) {
    const double sign = H->sign_rule ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) {
      const int zz = (int)s[i] * (int)s[(i + 1) % n];
      e += H->J * H->delta * zz;
      e += sign * H->J * (1 - zz) * ratio2(i, (i + 1) % n);
    }
  } else {
    const int sign = H->sign_rule ? -1 : 1;
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)s[i] * (int)s[(i + 1) % n]);
      e += -H->J * yy;
      e += H->J * sign * (1.0 + yy) * ratio2(i, (i + 1) % n);
    }
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)s[i] * (int)s[(i + 2) % n]);
      e += -H->J2 * yy;
      e += H->J2 * (1.0 + yy) * ratio2(i, (i + 2) % n);
    }
  }
  return e;
}

}  // namespace oracle_detail
