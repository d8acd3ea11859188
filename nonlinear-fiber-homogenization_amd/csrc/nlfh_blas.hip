// Vector kernels of the solver: stiffness scale, norms and means with deterministic two-level
// FP64 reductions, mean projection, precision conversion for iterative refinement, the PCG scalar
// recurrences, and the homogenized-stiffness (energy) kernel. gfx950.
#include "nlfh_kernels.h"

namespace nlfh {

// s = emin + (1 - emin) clip01(rho)^p   (nlfiber/linear.py:27-30, 137-141); optionally
// ds/drho = p (1 - emin) clip01(rho)^(p-1) * (1 inside [0, 1] including the bounds, 0 outside),
// the derivative the reference's clip01 gives.
__global__ void stiffness_scale_kernel(const double* __restrict__ rho, double* __restrict__ s64,
                                       float* __restrict__ s32, double* __restrict__ ds, int64_t n, double emin,
                                       double p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double r = rho[i];
    const double x = r > 1.0 ? 1.0 : (r < 0.0 ? 0.0 : r);
    const double v = emin + (1.0 - emin) * pow(x, p);
    if (s64) s64[i] = v;
    if (s32) s32[i] = (float)v;
    if (ds) ds[i] = (r > 1.0 || r < 0.0) ? 0.0 : p * (1.0 - emin) * pow(x, p - 1.0);
  }
}

// Copy (with conversion) of voxel arrays.
template <typename Ti, typename To>
__global__ void convert_kernel(const Ti* __restrict__ in, To* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (To)in[i];
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ out, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v;
}


// Partial sums over a chunk of a cell's voxels, per rhs (grid.y) and component:
//   mode 0: sum v^2 (all components)      -> part[rhs][chunk]
//   mode 1: sum v per component            -> part[(rhs*3 + c)][chunk]
//   mode 2: sum a * b (all components)     -> part[rhs][chunk]
template <typename T>
__global__ void __launch_bounds__(256) reduce_kernel(const VecArgs va, const T* __restrict__ a,
                                                     const T* __restrict__ b, double* __restrict__ part,
                                                     int mode) {
  __shared__ double red[4];
  const ChunkItem ch = va.chunks[blockIdx.x];
  if (va.done != nullptr && va.done[ch.cell]) return;
  const CellDesc cd = va.cells[ch.cell];
  const int rhs = blockIdx.y;
  const int64_t base = cd.off_field + (int64_t)rhs * 3 * cd.nvox;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x = (double)a[base + c * cd.nvox + i];
      if (mode == 0) acc[c] += x * x;
      else if (mode == 1) acc[c] += x;
      else acc[c] += x * (double)b[base + c * cd.nvox + i];
    }
  }
  if (mode == 1) {
    for (int c = 0; c < 3; ++c) {
      const double t = block_sum(acc[c], red);
      if (threadIdx.x == 0) part[((int64_t)rhs * 3 + c) * va.nchunks + blockIdx.x] = t;
    }
  } else {
    const double t = block_sum(acc[0] + acc[1] + acc[2], red);
    if (threadIdx.x == 0) part[(int64_t)rhs * va.nchunks + blockIdx.x] = t;
  }
}

// Gram partials of two field sets: part[I*m + J][chunk] = sum over the chunk's voxels and the 3
// components of a_I * b_J (a, b component-planar with m fields per cell).
__global__ void __launch_bounds__(256) gram_kernel(const VecArgs va, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double red[4];
  const ChunkItem ch = va.chunks[blockIdx.x];
  const CellDesc cd = va.cells[ch.cell];
  const int m = va.m;
  double acc[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) acc[k] = 0.0;
  for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
    for (int c = 0; c < 3; ++c) {
      double av[6], bv[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        const int64_t off = cd.off_field + ((int64_t)r * 3 + c) * cd.nvox + i;
        av[r] = r < m ? a[off] : 0.0;
        bv[r] = r < m ? b[off] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[r * 6 + q] += av[r] * bv[q];
    }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if (r < m && q < m) {  // uniform over the block
        const double t = block_sum(acc[r * 6 + q], red);
        if (threadIdx.x == 0) part[(int64_t)(r * m + q) * va.nchunks + blockIdx.x] = t;
      }
    }
}

// out_I = sum_J coef[cell][I][J] x_J per voxel and component (out may alias x: each thread reads all
// m inputs of its voxel before writing).
__global__ void __launch_bounds__(256) combine_kernel(const VecArgs va, const double* x, const double* coef,
                                                      double* out) {
  const ChunkItem ch = va.chunks[blockIdx.x];
  const CellDesc cd = va.cells[ch.cell];
  const int m = va.m;
  const double* A = coef + (int64_t)ch.cell * m * m;
  for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
    for (int c = 0; c < 3; ++c) {
      double xv[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) xv[r] = r < m ? x[cd.off_field + ((int64_t)r * 3 + c) * cd.nvox + i] : 0.0;
      for (int r = 0; r < m; ++r) {
        double y = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) y += (q < m ? A[r * m + q] : 0.0) * xv[q];
        out[cd.off_field + ((int64_t)r * 3 + c) * cd.nvox + i] = y;
      }
    }
  }
}

// Per-cell, per-row sums of partials in a fixed order: out[cell*out_stride + row] = sum over the
// cell's item range of part[row][item], times `scale`, or times 1/nvox of the cell when `cells` is
// given (means). One thread per (cell, row).
__global__ void sum_partials_kernel(const double* __restrict__ part, int nitems, const int* __restrict__ range,
                                    int ncells, int rows, double scale, const CellDesc* __restrict__ cells,
                                    int out_stride, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * rows) return;
  const int cell = t / rows, row = t - cell * rows;
  double acc = 0.0;
  for (int w = range[2 * cell]; w < range[2 * cell + 1]; ++w) acc += part[(int64_t)row * nitems + w];
  const double sc = cells != nullptr ? 1.0 / (double)cells[cell].nvox : scale;
  out[(int64_t)cell * out_stride + row] = acc * sc;
}

// v -= mean (per cell, rhs, component); mean given per (cell, rhs*3 + c).
template <typename T>
__global__ void sub_mean_kernel(const VecArgs va, T* __restrict__ v, const double* __restrict__ mean) {
  const ChunkItem ch = va.chunks[blockIdx.x];
  const CellDesc cd = va.cells[ch.cell];
  const int rhs = blockIdx.y;
  const int64_t base = cd.off_field + (int64_t)rhs * 3 * cd.nvox;
  for (int c = 0; c < 3; ++c) {
    const T mu = (T)mean[((int64_t)ch.cell * va.m + rhs) * 3 + c];
    for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) v[base + c * cd.nvox + i] -= mu;
  }
}

// Iterative refinement: r32 = r64 / scale ; d32 = 0   (scale per (cell, rhs)). With pdir, the search
// direction an inner solve left behind (held as scale_prev * p_K by ir_accumulate_kernel) is divided by the
// new scale too, which puts it in the units of the next inner system K d = r / ||r||.
__global__ void ir_scale_down_kernel(const VecArgs va, const double* __restrict__ r64, float* __restrict__ r32,
                                     float* __restrict__ d32, const double* __restrict__ scale,
                                     float* __restrict__ pdir) {
  const ChunkItem ch = va.chunks[blockIdx.x];
  if (va.done != nullptr && va.done[ch.cell]) return;
  const CellDesc cd = va.cells[ch.cell];
  const int rhs = blockIdx.y;
  const double sc = scale[ch.cell * va.m + rhs];
  const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
  const int64_t base = cd.off_field + (int64_t)rhs * 3 * cd.nvox;
  for (int c = 0; c < 3; ++c)
    for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
      const int64_t o = base + c * cd.nvox + i;
      r32[o] = (float)(r64[o] * inv);
      d32[o] = 0.0f;
      if (pdir != nullptr) pdir[o] = (float)((double)pdir[o] * inv);
    }
}

// dst = src for the cells with flag[cell] != 0 (m fields of the plan's layout)
template <typename T>
__global__ void __launch_bounds__(256) copy_flagged_kernel(const VecArgs va, const int* __restrict__ flag,
                                                           const T* __restrict__ src, T* __restrict__ dst) {
  const ChunkItem ch = va.chunks[blockIdx.x];
  if (!flag[ch.cell]) return;
  const CellDesc cd = va.cells[ch.cell];
  const int64_t base = cd.off_field + (int64_t)blockIdx.y * 3 * cd.nvox;
  for (int c = 0; c < 3; ++c)
    for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) dst[base + c * cd.nvox + i] = src[base + c * cd.nvox + i];
}

template <typename T>
hipError_t launch_copy_flagged(const VecArgs& va, const int* flag, const T* src, T* dst, hipStream_t st) {
  hipLaunchKernelGGL((copy_flagged_kernel<T>), dim3(va.nchunks, va.m), dim3(256), 0, st, va, flag, src, dst);
  return hipGetLastError();
}
template hipError_t launch_copy_flagged<float>(const VecArgs&, const int*, const float*, float*, hipStream_t);
template hipError_t launch_copy_flagged<double>(const VecArgs&, const int*, const double*, double*, hipStream_t);

// Iterative refinement: x64 += scale * d32. With p0/p1, the last search direction of the inner solve,
// p_K in P[K & 1] (K = it[cell], iteration k reads its direction from P[k & 1]), is kept in P[0] as
// scale * p_K, its value in the units of the FP64 residual.
__global__ void ir_accumulate_kernel(const VecArgs va, double* __restrict__ x64, const float* __restrict__ d32,
                                     const double* __restrict__ scale, float* p0, const float* p1,
                                     const int* __restrict__ it) {
  const ChunkItem ch = va.chunks[blockIdx.x];
  if (va.done != nullptr && va.done[ch.cell]) return;
  const CellDesc cd = va.cells[ch.cell];
  const int rhs = blockIdx.y;
  const double sc = scale[ch.cell * va.m + rhs];
  const int64_t base = cd.off_field + (int64_t)rhs * 3 * cd.nvox;
  const float* src = (p0 != nullptr && (it[ch.cell] & 1)) ? p1 : p0;
  for (int c = 0; c < 3; ++c)
    for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
      const int64_t o = base + c * cd.nvox + i;
      x64[o] += sc * (double)d32[o];
      if (p0 != nullptr) p0[o] = (float)(sc * (double)src[o]);
    }
}

// PCG scalar recurrences (nlfiber/linear.py:173-203), one block per cell, thread t = rhs.


// The scalar kernels below run one block per cell and one wavefront per right-hand side
// (blockDim = 64 m).

// After the first preconditioner application: bn = ||b|| + 1e-300 (from the chunk partials of
// b), rn (from the ||r||^2 partials of pass A), rz; decide `done` (nlfiber/linear.py:176-188).
__global__ void cg_init_kernel(CgState st, const double* part_a, int nplanes, const int* prange, const double* part_z,
                               int nz_items, const int* zrange, const double* part_b, int nchunks, const int* crange) {
  __shared__ int notconv;
  const int cell = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) notconv = 0;
  __syncthreads();
  if (t < st.m) {
    const int k = cell * st.m + t;
    const double bn = sqrt(wave_sum(part_b, nchunks, t, crange, cell)) + 1e-300;
    const double rn = sqrt(wave_sum(part_a, nplanes, t, prange, cell));
    const double rz = wave_sum(part_z, nz_items, t, zrange, cell);
    if (lane == 0) {
      st.bn[k] = bn;
      st.rn[k] = rn;
      st.rz[k] = rz;
      st.beta[k] = 0.0;
      if (rn > st.tol[cell] * bn) atomicAdd(&notconv, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st.it[cell] = 0;
    const int skip = st.skip != nullptr ? st.skip[cell] : 0;
    st.done[cell] = (skip || notconv == 0 || st.maxit[cell] <= 0) ? 1 : 0;
  }
}

// alpha = rz / pAp (guard pAp == 0)
__global__ void cg_alpha_kernel(CgState st, const double* part_mv, int nmv, const int* mrange) {
  const int cell = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (st.done[cell] || t >= st.m) return;
  const int k = cell * st.m + t;
  const double pAp = wave_sum(part_mv, nmv, t, mrange, cell);
  if (lane == 0) {
    if (st.alpha_prev != nullptr) st.alpha_prev[k] = st.alpha[k];
    st.alpha[k] = st.rz[k] / (pAp == 0.0 ? 1.0 : pAp);
  }
}

// beta = rz1 / rz (guard), rz = rz1, rn = ||r||, it += 1, done when every rhs has rn <= tol*bn or
// it reaches maxit.
__global__ void cg_beta_kernel(CgState st, const double* part_a, int nplanes, const int* prange, const double* part_z,
                               int nz_items, const int* zrange) {
  __shared__ int notconv;
  const int cell = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (st.done[cell]) return;
  if (threadIdx.x == 0) notconv = 0;
  __syncthreads();
  if (t < st.m) {
    const int k = cell * st.m + t;
    const double rn = sqrt(wave_sum(part_a, nplanes, t, prange, cell));
    const double rz1 = wave_sum(part_z, nz_items, t, zrange, cell);
    if (lane == 0) {
      const double rz = st.rz[k];
      st.beta[k] = rz1 / (rz == 0.0 ? 1.0 : rz);
      st.rz[k] = rz1;
      st.rn[k] = rn;
      if (rn > st.tol[cell] * st.bn[k]) atomicAdd(&notconv, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int it = st.it[cell] + 1;
    st.it[cell] = it;
    if (notconv == 0 || it >= st.maxit[cell]) st.done[cell] = 1;
  }
}

// Per-(cell, rhs) norms from chunk partials of ||v||^2: out = sqrt(sum) + eps. When `flag` is given
// (iterative refinement), flag[cell] = 1 when every rhs has out <= tol * bn (or `force`), and when
// `tol_in` is given it receives the relative tolerance of the cell's next inner solve:
// max(base, safety * min_rhs tol * bn / out), i.e. no tighter than the last refinement step needs.
__global__ void norms_kernel(const double* part, int nchunks, const int* crange, int m, int M, double eps,
                             double* out, const double* bn, double tol, int* flag, const int* force, double* tol_in,
                             double tol_in_base, double tol_in_safety) {
  __shared__ int notconv;
  __shared__ double need[8];
  const int cell = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) notconv = 0;
  __syncthreads();
  if (t < m) {
    const int k = cell * M + t;
    const double v = sqrt(wave_sum(part, nchunks, t, crange, cell)) + eps;
    if (lane == 0) {
      out[k] = v;
      if (bn != nullptr && v > tol * bn[k]) atomicAdd(&notconv, 1);
      need[t] = (bn != nullptr && v > 0.0) ? tol * bn[k] / v : 1.0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (flag != nullptr) flag[cell] = (notconv == 0 || (force != nullptr && force[cell])) ? 1 : 0;
    if (tol_in != nullptr) {
      double mn = 1.0;
      for (int i = 0; i < m; ++i) mn = need[i] < mn ? need[i] : mn;
      const double want = tol_in_safety * mn;
      tol_in[cell] = want > tol_in_base ? want : tol_in_base;
    }
  }
}

// Homogenized stiffness (nlfiber/linear.py:220-229) in the Walsh-mode form: per element and
// right-hand side J, v_J = (total constant strain e0 + 4E_J, bilinear modes, trilinear mode) and
// w_J = L v_J (L the modal element operator at s = 1); b_IJ = v_I . w_J equals
// b_e(E_I + u_I, E_J + u_J). The block accumulates s_e b_IJ; optionally writes b_IJ / N per
// element (= dC_IJ / ds_e, nlfiber/linear.py:10-13).

__global__ void __launch_bounds__(256) energy_kernel(const EnergyArgs a) {
  extern __shared__ double smem_d[];
  double* V = smem_d;  // [eb][m][18]
  __shared__ double red[4];
  const ChunkItem ch = a.chunks[blockIdx.x];
  const CellDesc cd = a.cells[ch.cell];
  const int m = a.m;
  const int t = threadIdx.x;
  const int el = t / m, J = t - el * m;
  const bool act = el < a.eb;
  const int nx = cd.nx, ny = cd.ny, nz = cd.nz;
  const int64_t nvox = cd.nvox;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const double l1 = a.lam16, m1 = a.mu16;
  const double third = 1.0 / 3.0;
  const double l2 = l1 * third, m2 = m1 * third, a2 = l2 + 3.0 * m2;
  const double kap = (l1 + 4.0 * m1) / 9.0;
  for (int64_t e0 = ch.begin; e0 < ch.end; e0 += a.eb) {
    const int64_t e = e0 + el;
    const bool valid = act && e < ch.end;
    double v[18], w[18];
    double se = 0.0;
    if (valid) {
      const int k = (int)(e % nz);
      const int j = (int)((e / nz) % ny);
      const int i = (int)(e / ((int64_t)nz * ny));
      const int i1 = (i + 1) % nx, j1 = (j + 1) % ny, k1 = (k + 1) % nz;
      const double* U = a.u + cd.off_field + (int64_t)J * 3 * nvox;
      double h[3][8];
      for (int c = 0; c < 3; ++c) {
        const double* Uc = U + c * nvox;
        double nd[8];
        nd[0] = Uc[((int64_t)i * ny + j) * nz + k];
        nd[1] = Uc[((int64_t)i1 * ny + j) * nz + k];
        nd[2] = Uc[((int64_t)i * ny + j1) * nz + k];
        nd[3] = Uc[((int64_t)i1 * ny + j1) * nz + k];
        nd[4] = Uc[((int64_t)i * ny + j) * nz + k1];
        nd[5] = Uc[((int64_t)i1 * ny + j) * nz + k1];
        nd[6] = Uc[((int64_t)i * ny + j1) * nz + k1];
        nd[7] = Uc[((int64_t)i1 * ny + j1) * nz + k1];
        double S[4], D[4];
        for (int q = 0; q < 4; ++q) { S[q] = nd[2 * q + 1] + nd[2 * q]; D[q] = nd[2 * q + 1] - nd[2 * q]; }
        double Y[2][4];
        for (int kz = 0; kz < 2; ++kz) {
          Y[kz][0] = S[2 * kz + 1] + S[2 * kz];
          Y[kz][2] = S[2 * kz + 1] - S[2 * kz];
          Y[kz][1] = D[2 * kz + 1] + D[2 * kz];
          Y[kz][3] = D[2 * kz + 1] - D[2 * kz];
        }
        for (int mm = 0; mm < 4; ++mm) { h[c][mm] = Y[1][mm] + Y[0][mm]; h[c][mm + 4] = Y[1][mm] - Y[0][mm]; }
      }
      const double* E = a.strain + J * 6;
      double eps[6];
      eps[0] = h[0][1] + 4.0 * E[0];
      eps[1] = h[1][2] + 4.0 * E[1];
      eps[2] = h[2][4] + 4.0 * E[2];
      eps[3] = h[1][4] + h[2][2] + 4.0 * E[3];
      eps[4] = h[0][4] + h[2][1] + 4.0 * E[4];
      eps[5] = h[0][2] + h[1][1] + 4.0 * E[5];
      const double tr = eps[0] + eps[1] + eps[2];
      for (int q = 0; q < 6; ++q) v[q] = eps[q];
      w[0] = l1 * tr + 2.0 * m1 * eps[0];
      w[1] = l1 * tr + 2.0 * m1 * eps[1];
      w[2] = l1 * tr + 2.0 * m1 * eps[2];
      w[3] = m1 * eps[3]; w[4] = m1 * eps[4]; w[5] = m1 * eps[5];
      // bilinear modes: (xy, xz, yz) x components
      v[6] = h[0][3]; v[7] = h[1][3]; v[8] = h[2][3];
      v[9] = h[0][5]; v[10] = h[1][5]; v[11] = h[2][5];
      v[12] = h[0][6]; v[13] = h[1][6]; v[14] = h[2][6];
      const double tt = h[2][3] + h[1][5] + h[0][6];
      w[6] = a2 * h[0][3] + l2 * h[2][6];
      w[7] = a2 * h[1][3] + l2 * h[2][5];
      w[8] = m2 * (tt + h[2][3]);
      w[9] = a2 * h[0][5] + l2 * h[1][6];
      w[10] = m2 * (tt + h[1][5]);
      w[11] = a2 * h[2][5] + l2 * h[1][3];
      w[12] = m2 * (tt + h[0][6]);
      w[13] = a2 * h[1][6] + l2 * h[0][5];
      w[14] = a2 * h[2][6] + l2 * h[0][3];
      for (int c = 0; c < 3; ++c) { v[15 + c] = h[c][7]; w[15 + c] = kap * h[c][7]; }
      se = a.s[cd.off_vox + e];
      for (int q = 0; q < 18; ++q) V[((int64_t)el * m + J) * 18 + q] = v[q];
    }
    __syncthreads();
    if (valid) {
      for (int I = 0; I < m; ++I) {
        const double* vi = V + ((int64_t)el * m + I) * 18;
        double b = 0.0;
        for (int q = 0; q < 18; ++q) b += vi[q] * w[q];
        acc[I] += se * b;
        if (a.dcds) a.dcds[cd.off_vox * m * m + ((int64_t)I * m + J) * nvox + e] = b / (double)nvox;
      }
    }
    __syncthreads();
  }
  // reduce acc[I] over the block for each (I, J): thread (el, J) holds row I entries
  for (int I = 0; I < m; ++I) {
    for (int J2 = 0; J2 < m; ++J2) {
      const double val = (act && J == J2) ? acc[I] : 0.0;
      const double tsum = block_sum(val, red);
      if (threadIdx.x == 0) a.part[((int64_t)I * m + J2) * a.nchunks + blockIdx.x] = tsum;
    }
  }
}

// Host launchers -------------------------------------------------------------------------------
static inline int nblocks(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_stiffness_scale(const double* rho, double* s64, float* s32, double* ds, int64_t n, double emin,
                                  double p, hipStream_t st) {
  hipLaunchKernelGGL(stiffness_scale_kernel, dim3(nblocks(n, 256)), dim3(256), 0, st, rho, s64, s32, ds, n, emin, p);
  return hipGetLastError();
}

hipError_t launch_d2f(const double* in, float* out, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nblocks(n, 256)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

hipError_t launch_f2d(const float* in, double* out, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nblocks(n, 256)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

// dinv[v] = 1 / (kdiag * sum of s over the 8 elements that share node v), per cell (periodic): the inverse of the
// diagonal of K(s), which is the same for the three components since every diagonal entry of Ke is kdiag
__global__ void jacobi_dinv_kernel(const CellDesc* __restrict__ cells, int ncells, const double* __restrict__ s,
                                   double kdiag, float* __restrict__ dinv, int64_t tot) {
  // grid-stride: the launch caps its grid at 65536 blocks
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < tot; v += (int64_t)gridDim.x * blockDim.x) {
  int lo = 0, hi = ncells - 1;  // the cell holding voxel v: the last with off_vox <= v
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cells[mid].off_vox <= v) lo = mid;
    else hi = mid - 1;
  }
  const CellDesc cd = cells[lo];
  const int64_t l = v - cd.off_vox;
  const int k = (int)(l % cd.nz), j = (int)((l / cd.nz) % cd.ny), i = (int)(l / ((int64_t)cd.nz * cd.ny));
  const double* sc = s + cd.off_vox;
  double t = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ii = (a & 1) ? (i == 0 ? cd.nx - 1 : i - 1) : i;
    const int jj = (a & 2) ? (j == 0 ? cd.ny - 1 : j - 1) : j;
    const int kk = (a & 4) ? (k == 0 ? cd.nz - 1 : k - 1) : k;
    t += sc[((int64_t)ii * cd.ny + jj) * cd.nz + kk];
  }
  dinv[v] = (float)(1.0 / (kdiag * t));
  }
}

hipError_t launch_jacobi_dinv(const CellDesc* cells, int ncells, const double* s, double kdiag, float* dinv,
                              int64_t tot, hipStream_t st) {
  hipLaunchKernelGGL(jacobi_dinv_kernel, dim3(nblocks(tot, 256)), dim3(256), 0, st, cells, ncells, s, kdiag, dinv, tot);
  return hipGetLastError();
}

hipError_t launch_copy_d(const double* in, double* out, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((convert_kernel<double, double>), dim3(nblocks(n, 256)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(T* out, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((fill_kernel<T>), dim3(nblocks(n, 256)), dim3(256), 0, st, out, n, v);
  return hipGetLastError();
}
template hipError_t launch_fill<double>(double*, int64_t, double, hipStream_t);
template hipError_t launch_fill<float>(float*, int64_t, float, hipStream_t);
template hipError_t launch_fill<int>(int*, int64_t, int, hipStream_t);

template <typename T>
hipError_t launch_reduce(const VecArgs& va, const T* a, const T* b, double* part, int mode, hipStream_t st) {
  hipLaunchKernelGGL((reduce_kernel<T>), dim3(va.nchunks, va.m), dim3(256), 0, st, va, a, b, part, mode);
  return hipGetLastError();
}
template hipError_t launch_reduce<double>(const VecArgs&, const double*, const double*, double*, int, hipStream_t);
template hipError_t launch_reduce<float>(const VecArgs&, const float*, const float*, double*, int, hipStream_t);

// Paired iterate steps of the FP32 direction iteration (x += alpha_{k-1} p_{k-1} + alpha_k p_k in pass C of
// odd k): a cell whose last iteration K was even (it = K + 1 odd) still owes alpha_K p_K, and p_K is P[0]
// (iteration k reads its direction from P[k & 1]); alpha still holds alpha_K.
template <typename T>
__global__ void __launch_bounds__(256) x_flush_even_kernel(const VecArgs va, T* x, const T* p0, const double* alpha,
                                                           const int* it) {
  const ChunkItem ch = va.chunks[blockIdx.x];
  if ((it[ch.cell] & 1) == 0) return;
  const CellDesc cd = va.cells[ch.cell];
  const int rhs = blockIdx.y;
  const T a = (T)alpha[ch.cell * va.m + rhs];
  const int64_t base = cd.off_field + (int64_t)rhs * 3 * cd.nvox;
  for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) x[base + c * cd.nvox + i] += a * p0[base + c * cd.nvox + i];
  }
}

template <typename T>
hipError_t launch_x_flush_even(const VecArgs& va, T* x, const T* p0, const double* alpha, const int* it,
                               hipStream_t st) {
  hipLaunchKernelGGL((x_flush_even_kernel<T>), dim3(va.nchunks, va.m), dim3(256), 0, st, va, x, p0, alpha, it);
  return hipGetLastError();
}
template hipError_t launch_x_flush_even<float>(const VecArgs&, float*, const float*, const double*, const int*,
                                               hipStream_t);

__global__ void set_small_kernel(const SmallParams v, int n, double* dd, int* di) {
  const int t = threadIdx.x;
  if (t < n) {
    if (dd != nullptr) dd[t] = v.d[t];
    if (di != nullptr) di[t] = v.i[t];
  }
}

hipError_t launch_set_small(const SmallParams& v, int n, double* dd, int* di, hipStream_t st) {
  hipLaunchKernelGGL(set_small_kernel, dim3(1), dim3(64), 0, st, v, n, dd, di);
  return hipGetLastError();
}

hipError_t launch_gram(const VecArgs& va, const double* a, const double* b, double* part, hipStream_t st) {
  hipLaunchKernelGGL(gram_kernel, dim3(va.nchunks), dim3(256), 0, st, va, a, b, part);
  return hipGetLastError();
}

hipError_t launch_combine(const VecArgs& va, const double* x, const double* coef, double* out, hipStream_t st) {
  hipLaunchKernelGGL(combine_kernel, dim3(va.nchunks), dim3(256), 0, st, va, x, coef, out);
  return hipGetLastError();
}

hipError_t launch_sum_partials(const double* part, int nitems, const int* range, int ncells, int rows,
                               double scale, const CellDesc* cells, int out_stride, double* out, hipStream_t st) {
  const int n = ncells * rows;
  hipLaunchKernelGGL(sum_partials_kernel, dim3((n + 127) / 128), dim3(128), 0, st, part, nitems, range, ncells, rows,
                     scale, cells, out_stride, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sub_mean(const VecArgs& va, T* v, const double* mean, hipStream_t st) {
  hipLaunchKernelGGL((sub_mean_kernel<T>), dim3(va.nchunks, va.m), dim3(256), 0, st, va, v, mean);
  return hipGetLastError();
}
template hipError_t launch_sub_mean<double>(const VecArgs&, double*, const double*, hipStream_t);

hipError_t launch_ir_down(const VecArgs& va, const double* r64, float* r32, float* d32, const double* scale,
                          float* pdir, hipStream_t st) {
  hipLaunchKernelGGL(ir_scale_down_kernel, dim3(va.nchunks, va.m), dim3(256), 0, st, va, r64, r32, d32, scale, pdir);
  return hipGetLastError();
}

hipError_t launch_ir_acc(const VecArgs& va, double* x64, const float* d32, const double* scale, float* p0,
                         const float* p1, const int* it, hipStream_t st) {
  hipLaunchKernelGGL(ir_accumulate_kernel, dim3(va.nchunks, va.m), dim3(256), 0, st, va, x64, d32, scale, p0, p1,
                     it);
  return hipGetLastError();
}

hipError_t launch_cg_init(const CgState& cs, int ncells, const double* pa, int np, const int* pr, const double* pz,
                          int nz, const int* zr, const double* pb, int nch, const int* cr, hipStream_t st) {
  hipLaunchKernelGGL(cg_init_kernel, dim3(ncells), dim3(64 * cs.m), 0, st, cs, pa, np, pr, pz, nz, zr, pb, nch, cr);
  return hipGetLastError();
}
hipError_t launch_cg_alpha(const CgState& cs, int ncells, const double* pmv, int nmv, const int* mr, hipStream_t st) {
  hipLaunchKernelGGL(cg_alpha_kernel, dim3(ncells), dim3(64 * cs.m), 0, st, cs, pmv, nmv, mr);
  return hipGetLastError();
}
hipError_t launch_cg_beta(const CgState& cs, int ncells, const double* pa, int np, const int* pr, const double* pz,
                          int nz, const int* zr, hipStream_t st) {
  hipLaunchKernelGGL(cg_beta_kernel, dim3(ncells), dim3(64 * cs.m), 0, st, cs, pa, np, pr, pz, nz, zr);
  return hipGetLastError();
}

hipError_t launch_norms(const double* part, int nchunks, const int* crange, int ncells, int m, int M, double eps,
                        double* out, const double* bn, double tol, int* flag, const int* force, double* tol_in,
                        double tol_in_base, double tol_in_safety, hipStream_t st) {
  hipLaunchKernelGGL(norms_kernel, dim3(ncells), dim3(64 * m), 0, st, part, nchunks, crange, m, M, eps, out, bn, tol, flag,
                     force, tol_in, tol_in_base, tol_in_safety);
  return hipGetLastError();
}

hipError_t launch_energy(const EnergyArgs& a, hipStream_t st) {
  const int threads = ((a.eb * a.m + 63) / 64) * 64;
  const size_t lds = (size_t)a.eb * a.m * 18 * sizeof(double);
  hipLaunchKernelGGL(energy_kernel, dim3(a.nchunks), dim3(threads), lds, st, a);
  return hipGetLastError();
}

}  // namespace nlfh
