// Vector-Jacobian products of the time-gated absorption Jacobian without
// forming J.
//
// For y(m, g) = sum_k Re(C_gk data(m, k)), data(m, k) = (D phi_k)(d_m, s_m)
// and adjoint fields phiA (sources D^T s_fpsf), the derivative with respect
// to nodal mua at fixed kappa is
//   dy(m,g)/dmua_v = -Re sum_k C'_gk phiA(., k, d_m)^T C_v phi(., k, s_m),
//   C_v(i, j) = int N_v N_i N_j  (element-local, 1/20, 1/60, 1/120 of vol),
// (dMesh.m:206-247 replaces C_v by the nodal product, i.e. C_v -> dxyz^3
// e_v e_v^T).  With ybar the output gradient and D_mk = sum_g ybar_mg C'_gk:
//   E(i, k, s) = sum_{m : s_m = s} D_mk phiA(i, k, d_m)             (k_vjp_weights)
//   Z(i, j)    = sum_k sum_s phi(j, k, s) E(i, k, s)
//   exact:  g_v = -Re sum_{e ∋ v} vol_e sum_{i,j in e} c(v,i,j) Z(i,j) (k_vjp_exact)
//   nodal:  g_v = -dxyz^3 Re Z(v, v)                                  (k_vjp_nodal)
#include "common.hpp"

namespace dogpup {

// field: complex128 solver layout, ncols columns (sources first), width nsp.
__device__ inline long long fidx(int k, int c, int v, int nn, int nsp, int ncb) {
  return (((long long)k * ncb + c / nsp) * nn + v) * nsp + c % nsp;
}

// E[v][k][s] (complex128).  One thread per (v, k, s); the links of source s
// are link_idx[link_ptr[s] .. link_ptr[s+1]) (link m = (source col, detector col)).
__global__ void k_vjp_weights(const double* __restrict__ field, const double* __restrict__ D,
                              const int* __restrict__ link, const int* __restrict__ link_ptr,
                              const int* __restrict__ link_idx, double* __restrict__ E, int nn, int nf,
                              int ncols, int nsp, int ns, int nm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nn * nf * ns) return;
  const int s = (int)(t % ns);
  const long long vk = t / ns;
  const int k = (int)(vk % nf);
  const int v = (int)(vk / nf);
  const int ncb = (ncols + nsp - 1) / nsp;
  double er = 0.0, ei = 0.0;
  for (int q = link_ptr[s]; q < link_ptr[s + 1]; ++q) {
    const int m = link_idx[q];
    const int dcol = link[2 * m + 1];
    const double dr = D[2 * ((long long)k * nm + m)], di = D[2 * ((long long)k * nm + m) + 1];
    const long long o = fidx(k, dcol, v, nn, nsp, ncb);
    const double br = field[2 * o], bi = field[2 * o + 1];
    er += dr * br - di * bi;
    ei += dr * bi + di * br;
  }
  E[2 * t] = er;
  E[2 * t + 1] = ei;
}

// Re Z(i, j) = Re sum_k sum_s phi(j, k, s) E(i, k, s)
__device__ inline double zre(const double* __restrict__ field, const double* __restrict__ E, int i, int j,
                             int nn, int nf, int nsp, int ncb, int ns) {
  double acc = 0.0;
  for (int k = 0; k < nf; ++k) {
    const double* e = E + 2 * (((long long)i * nf + k) * ns);
    for (int s = 0; s < ns; ++s) {
      const long long o = fidx(k, s, j, nn, nsp, ncb);
      acc += field[2 * o] * e[2 * s] - field[2 * o + 1] * e[2 * s + 1];
    }
  }
  return acc;
}

__global__ void k_vjp_nodal(const double* __restrict__ field, const double* __restrict__ E,
                            double* __restrict__ g, int nn, int nf, int ncols, int nsp, int ns, double dx3) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int ncb = (ncols + nsp - 1) / nsp;
  g[v] = -dx3 * zre(field, E, v, v, nn, nf, nsp, ncb, ns);
}

// One thread per element: its 16 Z(i, j) and the scatter to its 4 nodes.
__global__ void k_vjp_exact(const double* __restrict__ field, const double* __restrict__ E,
                            const double* __restrict__ node, const int* __restrict__ elem,
                            double* __restrict__ g, int nn, int ne, int nf, int ncols, int nsp, int ns) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int ncb = (ncols + nsp - 1) / nsp;
  int n[4];
  double x[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    n[a] = elem[4 * e + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[a][d] = node[3 * n[a] + d];
  }
  double B[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) B[d][c] = x[c + 1][d] - x[0][d];
  const double det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                     B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                     B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
  const double vol = fabs(det) / 6.0;
  double Z[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Z[a][b] = zre(field, E, n[a], n[b], nn, nf, nsp, ncb, ns);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double c;
        if (a == v && b == v) c = 1.0 / 20.0;
        else if (a == v || b == v || a == b) c = 1.0 / 60.0;
        else c = 1.0 / 120.0;
        s += c * Z[a][b];
      }
    atomicAdd(&g[n[v]], -vol * s);
  }
}

}  // namespace dogpup

using namespace dogpup;

extern "C" int dogpup_jacobian_vjp(const double* d_field, const double* d_D, const int* d_link,
                                   const int* d_link_ptr, const int* d_link_idx, const double* d_node,
                                   const int* d_elem, double* d_E, double* d_g, int nn, int ne, int nf,
                                   int ncols, int nsp, int ns, int nm, int exact, double dxyz3,
                                   void* stream) {
  DP_ARG(d_field && d_D && d_link && d_link_ptr && d_link_idx && d_E && d_g, "NULL argument");
  DP_ARG(nn > 0 && nf > 0 && ncols > 0 && nsp > 0 && ns > 0 && ns <= ncols && nm > 0, "bad dimensions");
  DP_ARG(!exact || (d_node && d_elem && ne > 0), "exact mode needs the mesh");
  hipStream_t st = as_stream(stream);
  const long long nt = (long long)nn * nf * ns;
  k_vjp_weights<<<div_up(nt, 256), 256, 0, st>>>(d_field, d_D, d_link, d_link_ptr, d_link_idx, d_E, nn, nf,
                                                 ncols, nsp, ns, nm);
  DP_LAUNCH_CHECK();
  if (exact) {
    DP_HIP(hipMemsetAsync(d_g, 0, (size_t)nn * 8, st));
    k_vjp_exact<<<div_up(ne, 128), 128, 0, st>>>(d_field, d_E, d_node, d_elem, d_g, nn, ne, nf, ncols, nsp, ns);
  } else {
    k_vjp_nodal<<<div_up(nn, 256), 256, 0, st>>>(d_field, d_E, d_g, nn, nf, ncols, nsp, ns, dxyz3);
  }
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}
