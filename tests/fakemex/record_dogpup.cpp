// A recording stand-in for the three mex-level entry points of libdogpup_hip,
// for CPU tests of the gateways: each call keeps its integer arguments and
// a checksum of every input array over the length its sizes imply, and
// writes y[i] = i into the output (2 doubles per complex element), or fails
// with a set code and message.
#include <cstring>
#include <string>

#include "dogpup.h"

namespace {
struct Record {
  int fn = 0;               // 1 smv, 2 FSAImv, 3 FSAIP_gen
  int ints[4] = {0, 0, 0, 0};
  double sums[6] = {0, 0, 0, 0, 0, 0};
  int calls = 0;
};
Record g_rec;
int g_fail = 0;
std::string g_msg = "";

template <typename T>
double sum(const T* p, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += (double)p[i] * (double)((i % 7) + 1);  // position-weighted
  return s;
}

int finish(double* out, size_t ndouble) {
  ++g_rec.calls;
  if (g_fail) return g_fail;
  for (size_t i = 0; i < ndouble; ++i) out[i] = (double)i;
  return DOGPUP_OK;
}
}  // namespace

extern "C" {

const char* dogpup_last_error(void) { return g_msg.c_str(); }

int dogpup_smv_mex(double* h_y, const double* h_x, const double* h_f, const double* h_A, const int* h_rPtrA,
                   const int* h_cPtrA, int nrows, int ncols, int nfreqs, int nnz) {
  const size_t nel = (size_t)nrows * ncols * nfreqs;
  g_rec.fn = 1;
  const int in[4] = {nrows, ncols, nfreqs, nnz};
  std::memcpy(g_rec.ints, in, sizeof(in));
  const double s[6] = {sum(h_x, 2 * nel), sum(h_f, (size_t)nfreqs), sum(h_A, 2 * (size_t)nnz),
                       sum(h_rPtrA, (size_t)nrows + 1), sum(h_cPtrA, (size_t)nnz), 0.0};
  std::memcpy(g_rec.sums, s, sizeof(s));
  return finish(h_y, 2 * nel);
}

int dogpup_fsaimv_mex(double* h_y, const double* h_x, const double* h_G, const int* h_rPtrG,
                      const int* h_cPtrG, int nrows, int ncols, int nfreqs, int nnzG) {
  const size_t nel = (size_t)nrows * ncols * nfreqs;
  g_rec.fn = 2;
  const int in[4] = {nrows, ncols, nfreqs, nnzG};
  std::memcpy(g_rec.ints, in, sizeof(in));
  const double s[6] = {sum(h_x, 2 * nel), sum(h_G, 2 * (size_t)nnzG * nfreqs), sum(h_rPtrG, (size_t)nrows + 1),
                       sum(h_cPtrG, (size_t)nnzG), 0.0, 0.0};
  std::memcpy(g_rec.sums, s, sizeof(s));
  return finish(h_y, 2 * nel);
}

int dogpup_fsaip_gen_mex(double* h_valG, const double* h_valA, const int* h_rPtrA, const int* h_cPtrA,
                         const int* h_rPtrG, const int* h_cPtrG, int nrows, int nnzA, int nnzG, int nfreqs) {
  g_rec.fn = 3;
  const int in[4] = {nrows, nnzA, nnzG, nfreqs};
  std::memcpy(g_rec.ints, in, sizeof(in));
  const double s[6] = {sum(h_valA, 2 * (size_t)nnzA), sum(h_rPtrA, (size_t)nrows + 1), sum(h_cPtrA, (size_t)nnzA),
                       sum(h_rPtrG, (size_t)nrows + 1), sum(h_cPtrG, (size_t)nnzG), 0.0};
  std::memcpy(g_rec.sums, s, sizeof(s));
  return finish(h_valG, 2 * (size_t)nnzG * nfreqs);
}

// ---- test controls ----
void rec_get(int* fn, int* ints, double* sums, int* calls) {
  *fn = g_rec.fn;
  std::memcpy(ints, g_rec.ints, sizeof(g_rec.ints));
  std::memcpy(sums, g_rec.sums, sizeof(g_rec.sums));
  *calls = g_rec.calls;
}
void rec_reset(void) { g_rec = Record(); }
void rec_fail(int code, const char* msg) {
  g_fail = code;
  g_msg = msg ? msg : "";
}

}  // extern "C"
