// C entry for the CPU tests of dogpup::rb_weights (complex values as
// interleaved doubles).
#include <complex>
#include "rb_weights.hpp"

extern "C" void rbw(int n, const double* G, const double* cq, const double* sj, const double* fj, double skr,
                    double ski, double fk, double* y) {
  using cd = std::complex<double>;
  dogpup::rb_weights(n, reinterpret_cast<const cd*>(G), reinterpret_cast<const cd*>(cq),
                     reinterpret_cast<const cd*>(sj), fj, cd(skr, ski), fk, reinterpret_cast<cd*>(y));
}
