// P1-tetrahedral FEM assembly of the complex frequency-domain diffusion
// system on the CSR pattern of the full symmetric A.
//
// Element integrals (behaviour of DOGPUP/dMesh.m:1033-1194, whose Keast
// 5-point rule is exact for these polynomials):
//   K_e(a,b) = vol * kbar * (grad l_a . grad l_b) + vol * sum_k mua_k c_abk,
//   kbar = mean of the 4 nodal kappa, c_aaa = 1/20, c_aab = 1/60, c_abk = 1/120;
//   M_e(a,b) = vol * (1 + delta_ab) / 20.
// Robin faces (the face list of dMesh.m:1129-1140, prepared by the host):
//   B_f(a,b) = area * sum_q (1/3) R(x_q) N_a(x_q) N_b(x_q) with the 3-point
//   rule at (2/3,1/6), (1/6,2/3), (1/6,1/6) (dMesh.m:1218-1235).
// Output: valA_re = K + B, valA_im = scale_im * M (scale_im = df/c).
//
// MI355X design: one pass computes the 16 element-matrix entries of every
// tetrahedron (coalesced element-major stores), a second pass gives every
// CSR slot one lane that sums its contribution list (host-prepared,
// slot-major), so the scatter needs no atomics and is deterministic.
#include "common.hpp"

namespace dogpup {

__global__ __launch_bounds__(256) void k_element(int ne, const double* __restrict__ node,
                                                 const int* __restrict__ elem,
                                                 const double* __restrict__ mua,
                                                 const double* __restrict__ kappa,
                                                 double* __restrict__ work) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int n[4];
  double x[4][3], mu[4], kb = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    n[a] = elem[4 * e + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[a][d] = node[3 * n[a] + d];
    mu[a] = mua[n[a]];
    kb += kappa[n[a]];
  }
  kb *= 0.25;
  // B = [x1-x0, x2-x0, x3-x0] as columns; grad l_a (a=1..3) = rows of B^-1
  double B[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) B[d][c] = x[c + 1][d] - x[0][d];
  const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
  const double c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
  const double c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
  const double det = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
  const double inv = 1.0 / det;
  double gr[4][3];
  // rows of B^-1 = adj(B)/det, adj(B)[i][j] = cofactor(B)[j][i]
  gr[1][0] = c00 * inv;
  gr[2][0] = c01 * inv;
  gr[3][0] = c02 * inv;
  gr[1][1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * inv;
  gr[2][1] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * inv;
  gr[3][1] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * inv;
  gr[1][2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * inv;
  gr[2][2] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * inv;
  gr[3][2] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * inv;
#pragma unroll
  for (int d = 0; d < 3; ++d) gr[0][d] = -(gr[1][d] + gr[2][d] + gr[3][d]);
  const double vol = fabs(det) / 6.0;
  const double musum = mu[0] + mu[1] + mu[2] + mu[3];
  double* w = work + 17LL * e;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double g = gr[a][0] * gr[b][0] + gr[a][1] * gr[b][1] + gr[a][2] * gr[b][2];
      double m;
      if (a == b) m = mu[a] / 20.0 + (musum - mu[a]) / 60.0;
      else m = (mu[a] + mu[b]) / 60.0 + (musum - mu[a] - mu[b]) / 120.0;
      w[4 * a + b] = vol * kb * g + vol * m;
    }
  w[16] = vol;
}

__global__ __launch_bounds__(256) void k_slots(int nnz, int ne, const double* __restrict__ node,
                                               const int* __restrict__ bface,
                                               const double* __restrict__ R,
                                               const int* __restrict__ slot_ptr,
                                               const int* __restrict__ slot_src,
                                               const double* __restrict__ work, double scale_im,
                                               double* __restrict__ vre, double* __restrict__ vim) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  double re = 0.0, mm = 0.0;
  const long long ecodes = 16LL * ne;
  for (int q = slot_ptr[s]; q < slot_ptr[s + 1]; ++q) {
    const long long c = slot_src[q];
    if (c < ecodes) {
      const long long e = c >> 4;
      const int ab = (int)(c & 15);
      re += work[17 * e + ab];
      const double vol = work[17 * e + 16];
      mm += ((ab >> 2) == (ab & 3)) ? vol / 10.0 : vol / 20.0;
    } else {
      const long long fc = c - ecodes;
      const long long f = fc / 9;
      const int ab = (int)(fc % 9), a = ab / 3, b = ab % 3;
      const int n0 = bface[3 * f], n1 = bface[3 * f + 1], n2 = bface[3 * f + 2];
      double u[3], v[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        u[d] = node[3 * n1 + d] - node[3 * n0 + d];
        v[d] = node[3 * n2 + d] - node[3 * n0 + d];
      }
      const double cx = u[1] * v[2] - u[2] * v[1];
      const double cy = u[2] * v[0] - u[0] * v[2];
      const double cz = u[0] * v[1] - u[1] * v[0];
      const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
      const double Rn[3] = {R[n0], R[n1], R[n2]};
      const double px[3] = {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0};
      const double py[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
      double acc = 0.0;
#pragma unroll
      for (int qp = 0; qp < 3; ++qp) {
        const double N[3] = {1.0 - px[qp] - py[qp], px[qp], py[qp]};
        const double Rq = Rn[0] * N[0] + Rn[1] * N[1] + Rn[2] * N[2];
        acc += (1.0 / 3.0) * Rq * N[a] * N[b];
      }
      re += area * acc;
    }
  }
  vre[s] = re;
  vim[s] = scale_im * mm;
}

}  // namespace dogpup

using namespace dogpup;

extern "C" int dogpup_assemble(int nn, int ne, int nbf, int nnz, const double* d_node,
                               const int* d_elem, const int* d_bface, const double* d_mua,
                               const double* d_kappa, const double* d_R, const int* d_slot_ptr,
                               const int* d_slot_src, double scale_im, double* d_work,
                               double* d_valA_re, double* d_valA_im, void* stream) {
  DP_ARG(nn > 0 && ne > 0 && nbf >= 0 && nnz > 0, "bad dimensions");
  DP_ARG(d_node && d_elem && d_mua && d_kappa && d_R && d_slot_ptr && d_slot_src && d_work &&
             d_valA_re && d_valA_im && (nbf == 0 || d_bface),
         "NULL argument");
  DP_ARG(16LL * ne + 9LL * nbf < 2147483647LL, "too many contributions for int32 codes");
  hipStream_t st = as_stream(stream);
  k_element<<<div_up(ne, 256), 256, 0, st>>>(ne, d_node, d_elem, d_mua, d_kappa, d_work);
  DP_LAUNCH_CHECK();
  k_slots<<<div_up(nnz, 256), 256, 0, st>>>(nnz, ne, d_node, d_bface, d_R, d_slot_ptr, d_slot_src,
                                            d_work, scale_im, d_valA_re, d_valA_im);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}
