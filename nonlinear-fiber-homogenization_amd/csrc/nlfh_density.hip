// Density map phi (symmetric octant) -> rho (full cell) on the GPU, and its vector-Jacobian product
// (nlfiber/density.py:23-91, nlfiber/symmetry.py:26-102):
//
//   f   = F mirror(sym_perm(phi))          periodic hat filter as a direct stencil (the kernel has
//                                          radius r voxels, so |offset| <= R = max nonzero offset)
//   u   = exact_symmetrize(f / std(f))     every voxel takes its orbit's value, computed once from the
//                                          sorted octant representative (bitwise equal over the orbit)
//   rho = sigmoid(beta (u - eta)),  mean(rho) = v   (eta: a grid-wide bracket search, 64 candidates
//                                                    per pass, then 3 Newton steps)
//
// The adjoint: threshold with the implicit d eta, the orbit average (self-adjoint), the std scaling,
// the same stencil (the hat kernel is even), and the octant collapse (sum of the 8 mirror images,
// averaged over the 6 axis permutations). Everything is float64 and every reduction has a fixed order.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "nlfh.h"
#include "nlfh_common.h"

namespace nlfh {

__device__ __forceinline__ int fold_index(int i, int n) { return i < n / 2 ? i : n - 1 - i; }

__device__ __forceinline__ double sig(double t) { return 0.5 * (1.0 + tanh(0.5 * t)); }

// axis permutations in itertools order; transpose(x, p)[i0, i1, i2] = x[j] with j[p[a]] = i_a
__device__ __forceinline__ void perm_index(int pidx, int i0, int i1, int i2, int& a, int& b, int& c) {
  int j[3];
  const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  j[P[pidx][0]] = i0;
  j[P[pidx][1]] = i1;
  j[P[pidx][2]] = i2;
  a = j[0];
  b = j[1];
  c = j[2];
}

// f = mirror(sym_perm(phi)) over rows of k fields
__global__ void dm_expand_kernel(const double* __restrict__ phi, int no, double* __restrict__ out, int k) {
  const int n = 2 * no;
  const int64_t N = (int64_t)n * n * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / N);
    const int64_t v = i - row * N;
    const int z = (int)(v % n), y = (int)((v / n) % n), x = (int)(v / ((int64_t)n * n));
    const int t0 = fold_index(x, n), t1 = fold_index(y, n), t2 = fold_index(z, n);
    const double* ph = phi + (int64_t)row * no * no * no;
    double acc = 0.0;
    for (int p = 0; p < 6; ++p) {
      int a, b, c;
      perm_index(p, t0, t1, t2, a, b, c);
      acc += ph[((int64_t)a * no + b) * no + c];
    }
    out[i] = acc / 6.0;
  }
}

// out = periodic stencil of in with weights w[(ox + R)][(oy + R)][(oz + R)] (offsets in [-R, R])
__global__ void dm_conv_kernel(const double* __restrict__ in, int n, const double* __restrict__ w, int R,
                               double* __restrict__ out, int k) {
  const int64_t N = (int64_t)n * n * n;
  const int W = 2 * R + 1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / N);
    const int64_t v = i - row * N;
    const int z = (int)(v % n), y = (int)((v / n) % n), x = (int)(v / ((int64_t)n * n));
    const double* src = in + row * N;
    double acc = 0.0;
    for (int ox = -R; ox <= R; ++ox) {
      const int xx = (x + ox + n) % n;
      for (int oy = -R; oy <= R; ++oy) {
        const int yy = (y + oy + n) % n;
        for (int oz = -R; oz <= R; ++oz) {
          const double wt = w[((ox + R) * W + (oy + R)) * W + (oz + R)];
          if (wt != 0.0) acc += wt * src[((int64_t)xx * n + yy) * n + (z + oz + n) % n];
        }
      }
    }
    out[i] = acc;
  }
}

// value of the orbit of voxel v of a full-cell field x (times scale): the 48-map average computed
// as the reference's exact_symmetrize does (fold: sequential mirror averages per axis; sym_perm over
// the fold; the sorted representative)
__device__ double orbit_value(const double* __restrict__ x, int n, int64_t v, double scale) {
  const int z = (int)(v % n), y = (int)((v / n) % n), xx = (int)(v / ((int64_t)n * n));
  int t[3] = {fold_index(xx, n), fold_index(y, n), fold_index(z, n)};
  // sort ascending (three entries)
  if (t[0] > t[1]) { const int s = t[0]; t[0] = t[1]; t[1] = s; }
  if (t[1] > t[2]) { const int s = t[1]; t[1] = t[2]; t[2] = s; }
  if (t[0] > t[1]) { const int s = t[0]; t[0] = t[1]; t[1] = s; }
  double acc = 0.0;
  for (int p = 0; p < 6; ++p) {
    int a, b, c;
    perm_index(p, t[0], t[1], t[2], a, b, c);
    const int A = n - 1 - a, B = n - 1 - b, Cc = n - 1 - c;
    auto at = [&](int i, int j, int l) { return x[((int64_t)i * n + j) * n + l] * scale; };
    const double f00 = 0.5 * (0.5 * (at(a, b, c) + at(A, b, c)) + 0.5 * (at(a, B, c) + at(A, B, c)));
    const double f01 = 0.5 * (0.5 * (at(a, b, Cc) + at(A, b, Cc)) + 0.5 * (at(a, B, Cc) + at(A, B, Cc)));
    acc += 0.5 * (f00 + f01);
  }
  return acc / 6.0;
}

// out = orbit values of in * scale (scale = 1 / (std + 1e-12) from stats, or 1)
__global__ void dm_symmetrize_kernel(const double* __restrict__ in, int n, const double* __restrict__ stats,
                                     double* __restrict__ out, int k) {
  const int64_t N = (int64_t)n * n * n;
  const double scale = stats ? 1.0 / (stats[1] + 1e-12) : 1.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / N);
    const int64_t v = i - row * N;
    out[i] = orbit_value(in + row * N, n, v, scale);
  }
}

// block-wide sum (blockDim 1024), result in every thread
__device__ double block_sum_all(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

__device__ double block_min_all(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = red[0];
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 1; i < nw; ++i) t = fmin(t, red[i]);
  return t;
}

// stats[0] = mean(f), stats[1] = std(f) (population), one block
__global__ void __launch_bounds__(1024) dm_moments_kernel(const double* __restrict__ f, int64_t N, double* stats) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) s += f[i];
  const double mean = block_sum_all(s, red) / (double)N;
  double q = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
    const double d = f[i] - mean;
    q += d * d;
  }
  const double var = block_sum_all(q, red) / (double)N;
  if (threadIdx.x == 0) {
    stats[0] = mean;
    stats[1] = sqrt(var);
  }
}

// eta with mean(sigmoid(beta (u - eta))) = v (nlfiber/density.py:43-63). The reference bisects 200
// times and takes 3 Newton steps; here each pass evaluates the volume at kEtaCand interior points of the
// bracket over the whole grid at once and keeps the sub-interval that contains the root, so 10 passes
// shrink the bracket below double resolution; then the same 3 Newton steps. stats: [2] eta, [3] lo,
// [4] hi.
constexpr int kEtaCand = 64;  // candidates per pass (blockDim of the partial kernel is kEtaCand * kEtaLanes)
// 16 lanes per candidate and 128 voxels per block: 8 sigmoids per thread, 256 blocks for a 32^3 cell (the
// search is latency-bound: FP64 tanh per voxel and candidate, so it needs the whole GPU's threads)
constexpr int kEtaLanes = 16;
constexpr int kEtaVox = 128;  // voxels per block of the partial kernel

// stats[3], stats[4] = min(u) - 40 / beta, max(u) + 40 / beta (one block)
__global__ void __launch_bounds__(1024) dm_bracket_kernel(const double* __restrict__ u, int64_t N, double beta,
                                                          double* stats) {
  __shared__ double red[16];
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
    mn = fmin(mn, u[i]);
    mx = fmax(mx, u[i]);
  }
  const double umin = block_min_all(mn, red);
  const double umax = -block_min_all(-mx, red);
  if (threadIdx.x == 0) {
    stats[3] = umin - 40.0 / beta;
    stats[4] = umax + 40.0 / beta;
  }
}

__device__ __forceinline__ double eta_candidate(double lo, double hi, int j) {
  return lo + (hi - lo) * (double)(j + 1) / (double)(kEtaCand + 1);
}

// part[block][j] = sum over the block's voxels of sigmoid(beta (u - eta_j))
__global__ void __launch_bounds__(kEtaCand* kEtaLanes) dm_eta_partial_kernel(const double* __restrict__ u, int64_t N,
                                                                             double beta, const double* stats,
                                                                             double* part) {
  __shared__ double acc[kEtaLanes][kEtaCand];
  const int j = threadIdx.x % kEtaCand, w = threadIdx.x / kEtaCand;
  const double eta = eta_candidate(stats[3], stats[4], j);
  const int64_t b0 = (int64_t)blockIdx.x * kEtaVox;
  const int64_t b1 = b0 + kEtaVox < N ? b0 + kEtaVox : N;
  double s = 0.0;
  for (int64_t i = b0 + w; i < b1; i += kEtaLanes) s += sig(beta * (u[i] - eta));
  acc[w][j] = s;
  __syncthreads();
  if (w == 0) {
    double t = 0.0;
    for (int q = 0; q < kEtaLanes; ++q) t += acc[q][j];
    part[(int64_t)blockIdx.x * kEtaCand + j] = t;
  }
}

// volumes of the candidates (fixed-order sums over the blocks: lane q of candidate j sums blocks q, q + 16,
// ..., then the 16 lane sums are added in order); the new bracket is the sub-interval where the volume
// crosses v (the volume decreases with eta). One block of kEtaCand x kEtaLanes threads.
__global__ void __launch_bounds__(kEtaCand* kEtaLanes) dm_eta_select_kernel(const double* part, int nblocks,
                                                                            int64_t N, double v, double* stats) {
  __shared__ int above[kEtaCand];
  __shared__ double acc[kEtaLanes][kEtaCand];
  const int j = threadIdx.x % kEtaCand, q = threadIdx.x / kEtaCand;
  const double lo = stats[3], hi = stats[4];
  double s = 0.0;
  for (int b = q; b < nblocks; b += kEtaLanes) s += part[(int64_t)b * kEtaCand + j];
  acc[q][j] = s;
  __syncthreads();
  if (q == 0) {
    double t = 0.0;
    for (int r = 0; r < kEtaLanes; ++r) t += acc[r][j];
    above[j] = t / (double)N > v ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int k = 0;  // number of candidates with volume above v: the root lies after candidate k - 1
    while (k < kEtaCand && above[k]) ++k;
    const double nlo = k == 0 ? lo : eta_candidate(lo, hi, k - 1);
    const double nhi = k == kEtaCand ? hi : eta_candidate(lo, hi, k);
    stats[3] = nlo;
    stats[4] = nhi;
    stats[2] = 0.5 * (nlo + nhi);
  }
}

// part[block][0..1] = sums of sigmoid and beta sigmoid (1 - sigmoid) at stats[2]
__global__ void __launch_bounds__(256) dm_newton_partial_kernel(const double* __restrict__ u, int64_t N, double beta,
                                                                const double* stats, double* part) {
  __shared__ double red[4];
  const double eta = stats[2];
  const int64_t b0 = (int64_t)blockIdx.x * kEtaVox;
  const int64_t b1 = b0 + kEtaVox < N ? b0 + kEtaVox : N;
  double s = 0.0, ds = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const double g = sig(beta * (u[i] - eta));
    s += g;
    ds += beta * g * (1.0 - g);
  }
  const double ts = block_sum(s, red);
  const double tds = block_sum(ds, red);
  if (threadIdx.x == 0) {
    part[2 * (int64_t)blockIdx.x] = ts;
    part[2 * (int64_t)blockIdx.x + 1] = tds;
  }
}

// one Newton step eta -= (mean s - v) / (-mean ds) (guard as the reference), one block
// (fixed-order sums: thread t adds blocks t, t + 256, ..., then block_sum)
__global__ void __launch_bounds__(256) dm_newton_step_kernel(const double* part, int nblocks, int64_t N, double v,
                                                             double* stats) {
  __shared__ double red[4];
  double s = 0.0, ds = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    s += part[2 * (int64_t)b];
    ds += part[2 * (int64_t)b + 1];
  }
  const double ts = block_sum(s, red);
  const double tds = block_sum(ds, red);
  if (threadIdx.x == 0) {
    const double dv = -tds / (double)N;
    stats[2] = stats[2] - (ts / (double)N - v) / (dv != 0.0 ? dv : -1.0);
  }
}

// rho = sigmoid(beta (u - eta)), d = rho (1 - rho)
__global__ void dm_sigmoid_kernel(const double* __restrict__ u, int64_t N, double beta, const double* stats,
                                  double* __restrict__ rho, double* __restrict__ d) {
  const double eta = stats[2];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const double s = sig(beta * (u[i] - eta));
    if (rho) rho[i] = s;
    if (d) d[i] = s * (1.0 - s);
  }
}

// sums[0] = sum d; sums[1 + row] = sum rbar_row d; sums[4 + row] = sum a_row b (a, b optional), one block
__global__ void __launch_bounds__(1024) dm_rowsums_kernel(const double* __restrict__ x, const double* __restrict__ d,
                                                          int64_t N, int k, double* sums, int base, int with_d) {
  __shared__ double red[16];
  if (with_d) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) s += d[i];
    const double t = block_sum_all(s, red);
    if (threadIdx.x == 0) sums[0] = t;
  }
  for (int row = 0; row < k; ++row) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) s += x[row * N + i] * d[i];
    const double t = block_sum_all(s, red);
    if (threadIdx.x == 0) sums[base + row] = t;
  }
}

// ubar = beta d (rbar - sums[1 + row] / sums[0])
__global__ void dm_threshold_vjp_kernel(const double* __restrict__ rbar, const double* __restrict__ d, int64_t N, int k,
                                        double beta, const double* sums, double* __restrict__ ubar) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / N);
    const int64_t v = i - row * N;
    ubar[i] = beta * d[v] * (rbar[i] - sums[1 + row] / sums[0]);
  }
}

// fbar = usbar / (sd + eps) - sums[4 + row] / (sd + eps)^2 (f - m) / (N sd)
__global__ void dm_std_vjp_kernel(const double* __restrict__ usbar, const double* __restrict__ f, int64_t N, int k,
                                  const double* stats, const double* sums, double* __restrict__ fbar) {
  const double m = stats[0], sd = stats[1], se = sd + 1e-12;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / N);
    const int64_t v = i - row * N;
    fbar[i] = usbar[i] / se - sums[4 + row] / (se * se) * (f[v] - m) / ((double)N * sd);
  }
}

// phibar = sym_perm(sum of the 8 mirror images of mbar) per octant entry
__global__ void dm_collapse_kernel(const double* __restrict__ mbar, int no, double* __restrict__ phibar, int k) {
  const int n = 2 * no;
  const int64_t N = (int64_t)n * n * n, No = (int64_t)no * no * no;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < No * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(i / No);
    const int64_t o = i - row * No;
    const int c0 = (int)(o % no), b0 = (int)((o / no) % no), a0 = (int)(o / ((int64_t)no * no));
    const double* src = mbar + row * N;
    double acc = 0.0;
    for (int p = 0; p < 6; ++p) {
      int a, b, c;
      perm_index(p, a0, b0, c0, a, b, c);
      const int A = n - 1 - a, B = n - 1 - b, Cc = n - 1 - c;
      auto at = [&](int x, int y, int z) { return src[((int64_t)x * n + y) * n + z]; };
      const double s0 = (at(a, b, c) + at(A, b, c)) + (at(a, B, c) + at(A, B, c));
      const double s1 = (at(a, b, Cc) + at(A, b, Cc)) + (at(a, B, Cc) + at(A, B, Cc));
      acc += s0 + s1;
    }
    phibar[i] = acc / 6.0;
  }
}

static inline int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

}  // namespace nlfh

using namespace nlfh;

struct nlfh_dmap {
  int device = 0;
  int no = 0, n = 0, R = 0;
  double r = 1.5, v = 0.3;
  int64_t N = 0;
  double* d_w = nullptr;      // stencil weights (2R+1)^3
  double* d_f = nullptr;      // filtered field
  double* d_u = nullptr;      // symmetrised, scaled field
  double* d_d = nullptr;      // rho (1 - rho)
  double* d_stats = nullptr;  // mean, std, eta, bracket lo, hi
  double* d_sums = nullptr;   // [0] sum d, [1..3] rbar . d, [4..6] usbar . f
  double* d_a = nullptr;      // 3 N work rows
  double* d_b = nullptr;      // 3 N work rows
  double* d_part = nullptr;   // eta search partials (blocks x candidates)
  int eta_blocks = 0;
  std::vector<void*> allocs;
};

namespace {
int dm_alloc(nlfh_dmap* dm, double** p, size_t n) {
  if (hipMalloc((void**)p, n * sizeof(double)) != hipSuccess) {
    set_error("hipMalloc failed in nlfh_dmap_create");
    return NLFH_ENOMEM;
  }
  dm->allocs.push_back(*p);
  return NLFH_OK;
}

// f, stats, u (and d, rho when beta > 0) for phi
int dm_forward(nlfh_dmap* dm, const double* phi, double beta, double* rho, hipStream_t st) {
  const int64_t N = dm->N;
  hipLaunchKernelGGL(dm_expand_kernel, dim3(grid_for(N)), dim3(256), 0, st, phi, dm->no, dm->d_a, 1);
  hipLaunchKernelGGL(dm_conv_kernel, dim3(grid_for(N)), dim3(256), 0, st, dm->d_a, dm->n, dm->d_w, dm->R, dm->d_f, 1);
  hipLaunchKernelGGL(dm_moments_kernel, dim3(1), dim3(1024), 0, st, dm->d_f, N, dm->d_stats);
  hipLaunchKernelGGL(dm_symmetrize_kernel, dim3(grid_for(N)), dim3(256), 0, st, dm->d_f, dm->n, dm->d_stats, dm->d_u,
                     1);
  if (beta > 0.0) {
    const int nb = dm->eta_blocks;
    hipLaunchKernelGGL(dm_bracket_kernel, dim3(1), dim3(1024), 0, st, dm->d_u, N, beta, dm->d_stats);
    for (int pass = 0; pass < 10; ++pass) {
      hipLaunchKernelGGL(dm_eta_partial_kernel, dim3(nb), dim3(kEtaCand * kEtaLanes), 0, st, dm->d_u, N, beta,
                         dm->d_stats, dm->d_part);
      hipLaunchKernelGGL(dm_eta_select_kernel, dim3(1), dim3(kEtaCand * kEtaLanes), 0, st, dm->d_part, nb, N,
                         dm->v, dm->d_stats);
    }
    for (int it = 0; it < 3; ++it) {
      hipLaunchKernelGGL(dm_newton_partial_kernel, dim3(nb), dim3(256), 0, st, dm->d_u, N, beta, dm->d_stats,
                         dm->d_part);
      hipLaunchKernelGGL(dm_newton_step_kernel, dim3(1), dim3(256), 0, st, dm->d_part, nb, N, dm->v, dm->d_stats);
    }
    hipLaunchKernelGGL(dm_sigmoid_kernel, dim3(grid_for(N)), dim3(256), 0, st, dm->d_u, N, beta, dm->d_stats, rho,
                       dm->d_d);
  }
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}
}  // namespace

extern "C" {

int nlfh_dmap_create(int n_o, double r, double v, nlfh_dmap** out) {
  if (!out || n_o < 1 || !(r > 0.0) || !(v > 0.0 && v < 1.0)) {
    set_error("nlfh_dmap_create: need n_o >= 1, r > 0 and 0 < v < 1");
    return NLFH_EINVAL;
  }
  *out = nullptr;
  const int n = 2 * n_o;
  // the reference's periodic hat kernel on the full grid (nlfiber/density.py:23-31), normalised
  std::vector<double> k((size_t)n * n * n, 0.0);
  double tot = 0.0;
  int R = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int l = 0; l < n; ++l) {
        const double di = std::min(i, n - i), dj = std::min(j, n - j), dl = std::min(l, n - l);
        const double val = std::max(0.0, r - std::sqrt(di * di + dj * dj + dl * dl));
        k[((size_t)i * n + j) * n + l] = val;
        tot += val;
        if (val > 0.0) R = std::max(R, (int)std::max(di, std::max(dj, dl)));
      }
  if (tot == 0.0) {
    k[0] = 1.0;
    tot = 1.0;
  }
  if (2 * R + 1 > n) {
    set_error("nlfh_dmap_create: the filter radius wraps around the cell (2 ceil(r) - 1 > 2 n_o)");
    return NLFH_EUNSUPPORTED;
  }
  const int W = 2 * R + 1;
  std::vector<double> w((size_t)W * W * W, 0.0);
  for (int ox = -R; ox <= R; ++ox)
    for (int oy = -R; oy <= R; ++oy)
      for (int oz = -R; oz <= R; ++oz)
        w[((size_t)(ox + R) * W + (oy + R)) * W + (oz + R)] =
            k[(((size_t)((ox + n) % n)) * n + (oy + n) % n) * n + (oz + n) % n] / tot;
  nlfh_dmap* dm = new nlfh_dmap();
  hipGetDevice(&dm->device);
  dm->no = n_o;
  dm->n = n;
  dm->R = R;
  dm->r = r;
  dm->v = v;
  dm->N = (int64_t)n * n * n;
  dm->eta_blocks = (int)((dm->N + kEtaVox - 1) / kEtaVox);
  int rc = NLFH_OK;
  if ((rc = dm_alloc(dm, &dm->d_w, w.size())) || (rc = dm_alloc(dm, &dm->d_f, dm->N)) ||
      (rc = dm_alloc(dm, &dm->d_u, dm->N)) || (rc = dm_alloc(dm, &dm->d_d, dm->N)) ||
      (rc = dm_alloc(dm, &dm->d_stats, 8)) || (rc = dm_alloc(dm, &dm->d_sums, 8)) ||
      (rc = dm_alloc(dm, &dm->d_a, 3 * dm->N)) || (rc = dm_alloc(dm, &dm->d_b, 3 * dm->N)) ||
      (rc = dm_alloc(dm, &dm->d_part, (size_t)dm->eta_blocks * kEtaCand))) {
    nlfh_dmap_destroy(dm);
    return rc;
  }
  if (hipMemcpy(dm->d_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("nlfh_dmap_create: weight upload failed");
    nlfh_dmap_destroy(dm);
    return NLFH_EHIP;
  }
  *out = dm;
  return NLFH_OK;
}

int nlfh_dmap_destroy(nlfh_dmap* dm) {
  if (!dm) return NLFH_OK;
  (void)hipSetDevice(dm->device);
  for (void* p : dm->allocs) (void)hipFree(p);
  delete dm;
  return NLFH_OK;
}

int nlfh_dmap_smooth(nlfh_dmap* dm, const double* phi, double* u_out, void* stream) {
  if (!dm || !phi || !u_out) {
    set_error("null argument");
    return NLFH_EINVAL;
  }
  HIP_TRY(hipSetDevice(dm->device));
  hipStream_t st = (hipStream_t)stream;
  int rc = dm_forward(dm, phi, 0.0, nullptr, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(u_out, dm->d_u, dm->N * sizeof(double), hipMemcpyDeviceToDevice, st));
  return NLFH_OK;
}

int nlfh_dmap_rho(nlfh_dmap* dm, const double* phi, double beta, double* rho_out, double* eta_out, void* stream) {
  if (!dm || !phi || !rho_out || !(beta > 0.0)) {
    set_error("nlfh_dmap_rho: null argument or beta <= 0");
    return NLFH_EINVAL;
  }
  HIP_TRY(hipSetDevice(dm->device));
  hipStream_t st = (hipStream_t)stream;
  int rc = dm_forward(dm, phi, beta, rho_out, st);
  if (rc) return rc;
  if (eta_out) {
    HIP_TRY(hipMemcpyAsync(eta_out, dm->d_stats + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return NLFH_OK;
}

int nlfh_dmap_rho_vjp(nlfh_dmap* dm, const double* phi, double beta, const double* rbar, int k, double* phibar_out,
                      void* stream) {
  if (!dm || !phi || !rbar || !phibar_out || k < 1 || k > 3 || !(beta > 0.0)) {
    set_error("nlfh_dmap_rho_vjp: null argument, beta <= 0 or k outside 1..3");
    return NLFH_EINVAL;
  }
  HIP_TRY(hipSetDevice(dm->device));
  hipStream_t st = (hipStream_t)stream;
  int rc = dm_forward(dm, phi, beta, nullptr, st);
  if (rc) return rc;
  const int64_t N = dm->N;
  const int g = grid_for(N * k);
  // threshold: ubar = beta d (rbar - (rbar . d) / sum d)              -> a
  hipLaunchKernelGGL(dm_rowsums_kernel, dim3(1), dim3(1024), 0, st, rbar, dm->d_d, N, k, dm->d_sums, 1, 1);
  hipLaunchKernelGGL(dm_threshold_vjp_kernel, dim3(g), dim3(256), 0, st, rbar, dm->d_d, N, k, beta, dm->d_sums,
                     dm->d_a);
  // orbit average (self-adjoint)                                       -> b
  hipLaunchKernelGGL(dm_symmetrize_kernel, dim3(g), dim3(256), 0, st, dm->d_a, dm->n, nullptr, dm->d_b, k);
  // std scaling: needs usbar . f per row                                -> a
  hipLaunchKernelGGL(dm_rowsums_kernel, dim3(1), dim3(1024), 0, st, dm->d_b, dm->d_f, N, k, dm->d_sums, 4, 0);
  hipLaunchKernelGGL(dm_std_vjp_kernel, dim3(g), dim3(256), 0, st, dm->d_b, dm->d_f, N, k, dm->d_stats, dm->d_sums,
                     dm->d_a);
  // filter (even kernel: the adjoint is the same stencil)             -> b
  hipLaunchKernelGGL(dm_conv_kernel, dim3(g), dim3(256), 0, st, dm->d_a, dm->n, dm->d_w, dm->R, dm->d_b, k);
  // mirror and sym_perm adjoints                                         -> phibar
  const int64_t No = (int64_t)dm->no * dm->no * dm->no;
  hipLaunchKernelGGL(dm_collapse_kernel, dim3(grid_for(No * k)), dim3(256), 0, st, dm->d_b, dm->no, phibar_out, k);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

}  // extern "C"
