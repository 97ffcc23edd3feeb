// C entry for the CPU tests of dogpup::rb_weights (complex values as
// interleaved doubles).
#include "rb_weights.hpp"

extern "C" void rbw(int n, const double* G, const double* cq, const double* sj, const double* fj, double skr,
                    double ski, double fk, double* y) {
  dogpup::rb_weights(n, G, cq, sj, fj, skr, ski, fk, y);
}
