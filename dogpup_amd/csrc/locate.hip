// Point location in a tetrahedral mesh and grid <-> mesh interpolation
// weights: the geometry under mesh2grid (DOGPUP/dMesh.m:286-380),
// mesh2mesh (:383-411), mesh_slice/mesh_slice3 (:414-517) and optode
// location (dOptode.m:77-97), all of which call MATLAB's
// triangulation/pointLocation on the host.
//
// Locator: the mesh's bounding box is cut into a uniform grid of bins; every
// tetrahedron is listed in each bin its (slightly padded) bounding box
// overlaps (count, exclusive scan, fill: a CSR of bins).  A query point reads
// its one bin and computes FP64 barycentric weights (Cramer's rule) for the
// tetrahedra listed there; it keeps the one whose smallest weight is largest
// (ties: smaller element id), and accepts it when that weight is >= -tol.
// With bins about 1.5x the mean element size a bin lists a few tens of
// tetrahedra, so a query costs a few thousand FP64 flops and one bin's worth
// of gathers; the whole locate is latency-bound at grid sizes and takes
// milliseconds.
//
// Grid -> mesh weights (g2m): the voxel grid is triangulated by the Kuhn
// (Freudenthal) split, six tetrahedra per cell around its main diagonal.  All
// eight vertices of a box cell lie on one sphere, so this is a Delaunay
// triangulation of the grid points, i.e. one of the triangulations the
// reference's delaunay(grid) may return.  A node with fractional cell
// coordinates f lies in the tetrahedron of the permutation p sorting f in
// descending order (stable on ties), with weights 1-f[p0], f[p0]-f[p1],
// f[p1]-f[p2], f[p2] on the vertices 0, e_p0, e_p0+e_p1, (1,1,1).
#include "common.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

namespace dogpup {

struct BinGrid {
  double lo[3];
  double inv[3];  // 1 / bin side
  int n[3];
};

__device__ inline int bin_coord(double x, double lo, double inv, int n) {
  int i = (int)floor((x - lo) * inv);
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// Bin range of element e's bounding box, padded by pad.
__device__ inline void elem_bins(const double* __restrict__ node, const int* __restrict__ elem, int e,
                                 const BinGrid& g, double pad, int b0[3], int b1[3]) {
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int q = 0; q < 4; ++q) {
    const int v = elem[4 * (long long)e + q];
    for (int a = 0; a < 3; ++a) {
      const double x = node[3 * (long long)v + a];
      lo[a] = fmin(lo[a], x);
      hi[a] = fmax(hi[a], x);
    }
  }
  for (int a = 0; a < 3; ++a) {
    b0[a] = bin_coord(lo[a] - pad, g.lo[a], g.inv[a], g.n[a]);
    b1[a] = bin_coord(hi[a] + pad, g.lo[a], g.inv[a], g.n[a]);
  }
}

__global__ void k_bin_count(const double* __restrict__ node, const int* __restrict__ elem, int ne, BinGrid g,
                            double pad, int* __restrict__ cnt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int b0[3], b1[3];
  elem_bins(node, elem, e, g, pad, b0, b1);
  for (int k = b0[2]; k <= b1[2]; ++k)
    for (int j = b0[1]; j <= b1[1]; ++j)
      for (int i = b0[0]; i <= b1[0]; ++i) atomicAdd(&cnt[i + g.n[0] * (j + (long long)g.n[1] * k)], 1);
}

__global__ void k_bin_fill(const double* __restrict__ node, const int* __restrict__ elem, int ne, BinGrid g,
                           double pad, int* __restrict__ cursor, int* __restrict__ items) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int b0[3], b1[3];
  elem_bins(node, elem, e, g, pad, b0, b1);
  for (int k = b0[2]; k <= b1[2]; ++k)
    for (int j = b0[1]; j <= b1[1]; ++j)
      for (int i = b0[0]; i <= b1[0]; ++i) {
        const int slot = atomicAdd(&cursor[i + g.n[0] * (j + (long long)g.n[1] * k)], 1);
        items[slot] = e;
      }
}

// In-place exclusive scan of cnt[0..n) into ptr[0..n], ptr[n] = total, by one
// workgroup of 1024 threads (each scans a contiguous chunk).  The total must
// fit in int (checked by the host from ptr[n] with a 64-bit total below).
__global__ void __launch_bounds__(1024) k_scan_one_block(const int* __restrict__ cnt, int* __restrict__ ptr,
                                                          long long n, long long* __restrict__ total64) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const long long chunk = (n + 1023) / 1024;
  const long long a = t * chunk, b = min(n, a + chunk);
  long long s = 0;
  for (long long i = a; i < b; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    long long v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - s;  // exclusive prefix of this chunk
  for (long long i = a; i < b; ++i) {
    ptr[i] = (int)run;
    run += cnt[i];
  }
  if (t == 1023) {
    ptr[n] = (int)part[1023];
    *total64 = part[1023];
  }
}

// Barycentric weights of p in tetrahedron (x0, x1, x2, x3) by Cramer's rule;
// returns false for a degenerate (zero-volume) tetrahedron.
__device__ inline bool bary_weights(const double* x0, const double* x1, const double* x2, const double* x3,
                                    const double* p, double w[4]) {
  double a[3], b[3], c[3], d[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = x1[k] - x0[k];
    b[k] = x2[k] - x0[k];
    c[k] = x3[k] - x0[k];
    d[k] = p[k] - x0[k];
  }
  auto det3 = [](const double* u, const double* v, const double* z) {
    return u[0] * (v[1] * z[2] - v[2] * z[1]) - u[1] * (v[0] * z[2] - v[2] * z[0]) +
           u[2] * (v[0] * z[1] - v[1] * z[0]);
  };
  const double det = det3(a, b, c);
  if (det == 0.0) return false;
  const double l1 = det3(d, b, c) / det;
  const double l2 = det3(a, d, c) / det;
  const double l3 = det3(a, b, d) / det;
  w[0] = 1.0 - l1 - l2 - l3;
  w[1] = l1;
  w[2] = l2;
  w[3] = l3;
  return isfinite(w[0]) && isfinite(l1) && isfinite(l2) && isfinite(l3);
}

__global__ void k_locate(const double* __restrict__ node, const int* __restrict__ elem, BinGrid g,
                         const int* __restrict__ bptr, const int* __restrict__ items, const double* __restrict__ pts,
                         long long npts, double tol, int* __restrict__ eid_out, double* __restrict__ w_out,
                         int* __restrict__ nodes_out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  double best = -DBL_MAX, bw[4] = {0, 0, 0, 0};
  int be = INT_MAX;
  bool inside_box = true;
  int bc[3];
  for (int a = 0; a < 3; ++a) {
    const double u = (p[a] - g.lo[a]) * g.inv[a];
    if (!(u >= -1e-9 && u <= g.n[a] + 1e-9)) inside_box = false;
    bc[a] = bin_coord(p[a], g.lo[a], g.inv[a], g.n[a]);
  }
  if (inside_box) {
    const long long b = bc[0] + g.n[0] * (bc[1] + (long long)g.n[1] * bc[2]);
    const int t0 = bptr[b], t1 = bptr[b + 1];
    for (int t = t0; t < t1; ++t) {
      const int e = items[t];
      const int v0 = elem[4 * (long long)e], v1 = elem[4 * (long long)e + 1];
      const int v2 = elem[4 * (long long)e + 2], v3 = elem[4 * (long long)e + 3];
      double x0[3], x1[3], x2[3], x3[3];
      for (int k = 0; k < 3; ++k) {
        x0[k] = node[3 * (long long)v0 + k];
        x1[k] = node[3 * (long long)v1 + k];
        x2[k] = node[3 * (long long)v2 + k];
        x3[k] = node[3 * (long long)v3 + k];
      }
      double w[4];
      if (!bary_weights(x0, x1, x2, x3, p, w)) continue;
      const double m = fmin(fmin(w[0], w[1]), fmin(w[2], w[3]));
      if (m > best || (m == best && e < be)) {
        best = m;
        be = e;
        for (int q = 0; q < 4; ++q) bw[q] = w[q];
      }
    }
  }
  const bool found = be != INT_MAX && best >= -tol;
  eid_out[i] = found ? be : -1;
  for (int q = 0; q < 4; ++q) {
    w_out[4 * i + q] = found ? bw[q] : NAN;
    if (nodes_out) nodes_out[4 * i + q] = found ? elem[4 * (long long)be + q] : -1;
  }
}

// Grid -> mesh weights of every node (Kuhn split of the grid, in-mesh filter
// and row normalisation as dMesh.m:331-373).  status[i]: 0 weights written,
// 1 no in-mesh vertex (nearest-voxel fallback needed), 2 outside the grid.
struct GridDesc {
  double o[3], h[3];
  int n[3];
};

__global__ void k_grid_to_mesh(const double* __restrict__ node, long long nn, GridDesc g,
                               const unsigned char* __restrict__ inmesh, double eps, int* __restrict__ idx,
                               double* __restrict__ w, int* __restrict__ status) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  int c[3];
  double f[3];
  bool outside = false;
  for (int a = 0; a < 3; ++a) {
    const double q = (node[3 * i + a] - g.o[a]) / g.h[a];
    if (q < -eps || q > (g.n[a] - 1) + eps) outside = true;
    int ci = (int)floor(q);
    ci = ci < 0 ? 0 : (ci > g.n[a] - 2 ? g.n[a] - 2 : ci);
    c[a] = ci;
    f[a] = fmin(1.0, fmax(0.0, q - ci));
  }
  for (int q = 0; q < 4; ++q) {
    idx[4 * i + q] = -1;
    w[4 * i + q] = 0.0;
  }
  if (outside) {
    status[i] = 2;
    return;
  }
  // stable descending order of the fractional coordinates
  int p[3] = {0, 1, 2};
  for (int s = 1; s < 3; ++s)
    for (int t = s; t > 0 && f[p[t]] > f[p[t - 1]]; --t) {
      const int tmp = p[t];
      p[t] = p[t - 1];
      p[t - 1] = tmp;
    }
  const double wt[4] = {1.0 - f[p[0]], f[p[0]] - f[p[1]], f[p[1]] - f[p[2]], f[p[2]]};
  int off[3] = {0, 0, 0};
  const long long sy = g.n[0], sz = (long long)g.n[0] * g.n[1];
  int kept = 0;
  double sum = 0.0;
  for (int q = 0; q < 4; ++q) {
    if (q > 0) off[p[q - 1]] = 1;
    const long long v = (c[0] + off[0]) + sy * (c[1] + off[1]) + sz * (c[2] + off[2]);
    if (inmesh[v]) {
      idx[4 * i + kept] = (int)v;
      w[4 * i + kept] = wt[q];
      sum += wt[q];
      ++kept;
    }
  }
  if (kept == 0) {
    status[i] = 1;
    return;
  }
  const double s = sum == 0.0 ? 1.0 : sum;
  for (int q = 0; q < kept; ++q) w[4 * i + q] /= s;
  status[i] = 0;
}

// Nearest in-mesh voxel (ties: smaller voxel index) for the listed nodes.
__global__ void k_nearest_voxel(const double* __restrict__ node, const int* __restrict__ list, int nlist, GridDesc g,
                                const int* __restrict__ vox, int nvox, int* __restrict__ idx,
                                double* __restrict__ w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const long long i = list[t];
  const double px = node[3 * i], py = node[3 * i + 1], pz = node[3 * i + 2];
  double best = DBL_MAX;
  int bv = INT_MAX;
  const int nx = g.n[0], ny = g.n[1];
  for (int k = 0; k < nvox; ++k) {
    const int v = vox[k];
    const int ix = v % nx, iy = (v / nx) % ny, iz = v / (nx * ny);
    const double dx = g.o[0] + ix * g.h[0] - px;
    const double dy = g.o[1] + iy * g.h[1] - py;
    const double dz = g.o[2] + iz * g.h[2] - pz;
    const double d = dx * dx + dy * dy + dz * dz;
    if (d < best || (d == best && v < bv)) {
      best = d;
      bv = v;
    }
  }
  idx[4 * i] = bv;
  w[4 * i] = 1.0;
}

// out[g][r][c] = sum_q w[r][q] * in[g][idx[r][q]][c] over the entries with
// idx >= 0; rows of width doubles, handled as double2 pairs.
__global__ void k_ell4_apply(const double* __restrict__ in, double* __restrict__ out, const int* __restrict__ idx,
                             const double* __restrict__ w, long long nout, long long nin, int width) {
  const int half = width >> 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout * half) return;
  const long long r = t / half;
  const int c = (int)(t - r * half);
  const double* src = in + (long long)blockIdx.y * nin * width;
  double* dst = out + (long long)blockIdx.y * nout * width;
  double2 acc = make_double2(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = idx[4 * r + q];
    if (j >= 0) {
      const double wq = w[4 * r + q];
      const double2 v = reinterpret_cast<const double2*>(src + (long long)j * width)[c];
      acc.x += wq * v.x;
      acc.y += wq * v.y;
    }
  }
  reinterpret_cast<double2*>(dst + r * width)[c] = acc;
}

__global__ void k_ell4_apply_odd(const double* __restrict__ in, double* __restrict__ out, const int* __restrict__ idx,
                                 const double* __restrict__ w, long long nout, long long nin, int width) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout * width) return;
  const long long r = t / width;
  const int c = (int)(t - r * width);
  const double* src = in + (long long)blockIdx.y * nin * width;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = idx[4 * r + q];
    if (j >= 0) acc += w[4 * r + q] * src[(long long)j * width + c];
  }
  out[(long long)blockIdx.y * nout * width + r * width + c] = acc;
}

}  // namespace dogpup

using namespace dogpup;

struct dogpup_locator {
  int nn = 0, ne = 0;
  double* node = nullptr;
  int* elem = nullptr;
  int* bptr = nullptr;
  int* items = nullptr;
  BinGrid g{};
  long long nbins = 0, nitems = 0;
};

static void locator_free(dogpup_locator* h) {
  if (!h) return;
  (void)hipFree(h->node);
  (void)hipFree(h->elem);
  (void)hipFree(h->bptr);
  (void)hipFree(h->items);
  delete h;
}

extern "C" int dogpup_locator_create(dogpup_locator_t* out, const double* d_node, int nn, const int* d_elem, int ne,
                                     void* stream) {
  DP_ARG(out && d_node && d_elem, "locator_create: null argument");
  DP_ARG(nn >= 4 && ne >= 1, "locator_create: need nn >= 4 and ne >= 1 (got %d, %d)", nn, ne);
  *out = nullptr;
  hipStream_t st = as_stream(stream);
  std::vector<double> hn(3 * (size_t)nn);
  DP_HIP(hipMemcpyAsync(hn.data(), d_node, hn.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < nn; ++i)
    for (int a = 0; a < 3; ++a) {
      const double x = hn[3 * (size_t)i + a];
      DP_ARG(std::isfinite(x), "locator_create: node %d has a non-finite coordinate", i);
      lo[a] = std::min(lo[a], x);
      hi[a] = std::max(hi[a], x);
    }
  double ext[3], vol = 1.0, diag = 0.0;
  for (int a = 0; a < 3; ++a) {
    ext[a] = hi[a] - lo[a];
    diag += ext[a] * ext[a];
  }
  diag = std::sqrt(diag);
  DP_ARG(diag > 0, "locator_create: all nodes coincide");
  for (int a = 0; a < 3; ++a) vol *= std::max(ext[a], 1e-6 * diag);
  // bin side ~ 1.5 x the cube root of the volume per element, at most 2^24 bins
  double side = 1.5 * std::cbrt(vol / ne);
  auto count_bins = [&](double s) {
    long long n = 1;
    for (int a = 0; a < 3; ++a) n *= std::max(1LL, (long long)std::ceil(ext[a] / s));
    return n;
  };
  while (count_bins(side) > (1LL << 24)) side *= 1.25;
  auto* h = new dogpup_locator();
  h->nn = nn;
  h->ne = ne;
  for (int a = 0; a < 3; ++a) {
    h->g.n[a] = (int)std::max(1LL, (long long)std::ceil(ext[a] / side));
    h->g.lo[a] = lo[a];
    h->g.inv[a] = ext[a] > 0 ? h->g.n[a] / ext[a] : 0.0;
  }
  h->nbins = (long long)h->g.n[0] * h->g.n[1] * h->g.n[2];
  const double pad = 1e-9 * diag;
  int rc = DOGPUP_OK;
  int* cnt = nullptr;
  long long* tot = nullptr;
  auto bail = [&](int code) {
    (void)hipFree(cnt);
    (void)hipFree(tot);
    locator_free(h);
    return code;
  };
#define LOC_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return bail(fail(DOGPUP_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(e_))); \
  } while (0)
  LOC_HIP(hipMalloc(&h->node, 3 * sizeof(double) * (size_t)nn));
  LOC_HIP(hipMalloc(&h->elem, 4 * sizeof(int) * (size_t)ne));
  LOC_HIP(hipMemcpyAsync(h->node, d_node, 3 * sizeof(double) * (size_t)nn, hipMemcpyDeviceToDevice, st));
  LOC_HIP(hipMemcpyAsync(h->elem, d_elem, 4 * sizeof(int) * (size_t)ne, hipMemcpyDeviceToDevice, st));
  LOC_HIP(hipMalloc(&cnt, sizeof(int) * (size_t)h->nbins));
  LOC_HIP(hipMalloc(&h->bptr, sizeof(int) * (size_t)(h->nbins + 1)));
  LOC_HIP(hipMalloc(&tot, sizeof(long long)));
  LOC_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)h->nbins, st));
  k_bin_count<<<div_up(ne, 256), 256, 0, st>>>(h->node, h->elem, ne, h->g, pad, cnt);
  LOC_HIP(hipGetLastError());
  k_scan_one_block<<<1, 1024, 0, st>>>(cnt, h->bptr, h->nbins, tot);
  LOC_HIP(hipGetLastError());
  long long total = 0;
  LOC_HIP(hipMemcpyAsync(&total, tot, sizeof(long long), hipMemcpyDeviceToHost, st));
  LOC_HIP(hipStreamSynchronize(st));
  if (total >= INT_MAX)
    return bail(fail(DOGPUP_ERR_UNSUPPORTED, "locator_create: %lld element-bin pairs exceed int32", total));
  h->nitems = total;
  LOC_HIP(hipMalloc(&h->items, sizeof(int) * (size_t)std::max(1LL, total)));
  LOC_HIP(hipMemcpyAsync(cnt, h->bptr, sizeof(int) * (size_t)h->nbins, hipMemcpyDeviceToDevice, st));
  k_bin_fill<<<div_up(ne, 256), 256, 0, st>>>(h->node, h->elem, ne, h->g, pad, cnt, h->items);
  LOC_HIP(hipGetLastError());
  LOC_HIP(hipStreamSynchronize(st));
#undef LOC_HIP
  (void)hipFree(cnt);
  (void)hipFree(tot);
  *out = h;
  return rc;
}

extern "C" int dogpup_locator_destroy(dogpup_locator_t h) {
  locator_free(h);
  return DOGPUP_OK;
}

extern "C" int dogpup_locator_stats(dogpup_locator_t h, long long* h_out) {
  DP_ARG(h && h_out, "locator_stats: null argument");
  h_out[0] = h->g.n[0];
  h_out[1] = h->g.n[1];
  h_out[2] = h->g.n[2];
  h_out[3] = h->nitems;
  return DOGPUP_OK;
}

extern "C" int dogpup_locator_find(dogpup_locator_t h, const double* d_pts, long long npts, double tol, int* d_eid,
                                   double* d_w, int* d_nodes, void* stream) {
  DP_ARG(h && d_eid && d_w, "locator_find: null argument");
  DP_ARG(npts >= 0, "locator_find: npts < 0");
  if (npts == 0) return DOGPUP_OK;
  DP_ARG(d_pts, "locator_find: null points");
  hipStream_t st = as_stream(stream);
  k_locate<<<div_up(npts, 128), 128, 0, st>>>(h->node, h->elem, h->g, h->bptr, h->items, d_pts, npts, tol, d_eid,
                                              d_w, d_nodes);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}

extern "C" int dogpup_grid_to_mesh(const double* d_node, int nn, const double* h_origin, const double* h_spacing,
                                   const int* h_dims, const unsigned char* d_inmesh, int* d_idx, double* d_w,
                                   int* h_status_counts, void* stream) {
  DP_ARG(d_node && h_origin && h_spacing && h_dims && d_inmesh && d_idx && d_w, "grid_to_mesh: null argument");
  DP_ARG(nn >= 1, "grid_to_mesh: nn < 1");
  GridDesc g;
  long long nv = 1;
  for (int a = 0; a < 3; ++a) {
    DP_ARG(h_dims[a] >= 2, "grid_to_mesh: every grid axis needs >= 2 points (axis %d has %d)", a, h_dims[a]);
    DP_ARG(h_spacing[a] > 0, "grid_to_mesh: spacing must be positive");
    g.o[a] = h_origin[a];
    g.h[a] = h_spacing[a];
    g.n[a] = h_dims[a];
    nv *= h_dims[a];
  }
  DP_ARG(nv < INT_MAX, "grid_to_mesh: %lld grid points exceed int32", nv);
  hipStream_t st = as_stream(stream);
  int* status = nullptr;
  DP_HIP(hipMalloc(&status, sizeof(int) * (size_t)nn));
  k_grid_to_mesh<<<div_up(nn, 256), 256, 0, st>>>(d_node, nn, g, d_inmesh, 1e-9, d_idx, d_w, status);
  hipError_t le = hipGetLastError();
  std::vector<int> hs(nn);
  std::vector<unsigned char> hin(nv);
  if (le == hipSuccess) le = hipMemcpyAsync(hs.data(), status, sizeof(int) * (size_t)nn, hipMemcpyDeviceToHost, st);
  if (le == hipSuccess) le = hipMemcpyAsync(hin.data(), d_inmesh, (size_t)nv, hipMemcpyDeviceToHost, st);
  if (le == hipSuccess) le = hipStreamSynchronize(st);
  (void)hipFree(status);
  if (le != hipSuccess) return fail(DOGPUP_ERR_HIP, "grid_to_mesh: %s", hipGetErrorString(le));
  std::vector<int> fb, vox;
  int nout = 0;
  for (int i = 0; i < nn; ++i) {
    if (hs[i] == 1) fb.push_back(i);
    if (hs[i] == 2) ++nout;
  }
  if (h_status_counts) {
    h_status_counts[0] = (int)fb.size();
    h_status_counts[1] = nout;
  }
  if (nout > 0)
    return fail(DOGPUP_ERR_ARG, "grid_to_mesh: %d nodes lie outside the voxel grid (the grid must cover the mesh)",
                nout);
  if (fb.empty()) return DOGPUP_OK;
  for (long long v = 0; v < nv; ++v)
    if (hin[v]) vox.push_back((int)v);
  if (vox.empty()) return fail(DOGPUP_ERR_ARG, "grid_to_mesh: no grid point lies inside the mesh");
  int *dl = nullptr, *dv = nullptr;
  hipError_t e = hipMalloc(&dl, sizeof(int) * fb.size());
  if (e == hipSuccess) e = hipMalloc(&dv, sizeof(int) * vox.size());
  if (e == hipSuccess) e = hipMemcpyAsync(dl, fb.data(), sizeof(int) * fb.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dv, vox.data(), sizeof(int) * vox.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    k_nearest_voxel<<<div_up((long long)fb.size(), 64), 64, 0, st>>>(d_node, dl, (int)fb.size(), g, dv,
                                                                     (int)vox.size(), d_idx, d_w);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // the host vectors are the copies' sources
  (void)hipFree(dl);
  (void)hipFree(dv);
  if (e != hipSuccess) return fail(DOGPUP_ERR_HIP, "grid_to_mesh: %s", hipGetErrorString(e));
  return DOGPUP_OK;
}

extern "C" int dogpup_ell4_apply(const double* d_in, double* d_out, const int* d_idx, const double* d_w,
                                 long long nrows_out, long long nrows_in, int width, int ngroups, void* stream) {
  DP_ARG(d_in && d_out && d_idx && d_w, "ell4_apply: null argument");
  DP_ARG(nrows_out >= 0 && nrows_in >= 0 && width >= 1 && ngroups >= 1 && ngroups <= 65535,
         "ell4_apply: bad sizes (nrows_out %lld, nrows_in %lld, width %d, ngroups %d)", nrows_out, nrows_in, width,
         ngroups);
  if (nrows_out == 0) return DOGPUP_OK;
  hipStream_t st = as_stream(stream);
  const bool pairs = (width % 2 == 0) && (reinterpret_cast<uintptr_t>(d_in) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(d_out) % 16 == 0);
  const long long work = pairs ? nrows_out * (width / 2) : nrows_out * width;
  dim3 grid(div_up(work, 256), ngroups);
  if (pairs)
    k_ell4_apply<<<grid, 256, 0, st>>>(d_in, d_out, d_idx, d_w, nrows_out, nrows_in, width);
  else
    k_ell4_apply_odd<<<grid, 256, 0, st>>>(d_in, d_out, d_idx, d_w, nrows_out, nrows_in, width);
  DP_LAUNCH_CHECK();
  return DOGPUP_OK;
}
