// Shared definitions of the gfx950 homogenization library: cell descriptors, work items,
// error handling and small device helpers.
//
// Field layout in HBM (all kernels): component-planar ("SoA") per cell and right-hand side,
//   field[cell][rhs][comp][x][y][z], z fastest,
// so a cell's block of m right-hand sides is contiguous and every z-row is a coalesced stream.
// The reference layout (m, nx, ny, nz, 3) (nlfiber/linear.py:80-92) is the same memory seen through
// the permutation (0, 2, 3, 4, 1); the Python layer exposes that view without a copy.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#define NLFH_MAX_CELLS 4096

namespace nlfh {

// One periodic cell of a (possibly batched) plan.
struct CellDesc {
  int nx, ny, nz;
  int nzh;            // nz / 2 + 1 (half spectrum along z)
  int shape_id;       // index of the distinct grid shape (symbol, twiddle tables)
  int pad_;
  int64_t nvox;       // nx * ny * nz
  int64_t off_vox;    // offset of the cell in voxel arrays (s, rho)
  int64_t off_field;  // offset of the cell in field arrays, in values: sum of m*3*nvox of earlier cells
  int64_t off_spec;   // offset of the cell in spectrum arrays, in complex values
  int64_t off_sym;    // offset of this shape's symbol (6 reals per frequency)
  int64_t nspec;      // nx * ny * nzh
};

// Work item of the matvec: a y-tile of TY node rows and an x-chunk of XC planes of one cell.
struct MvItem {
  int cell, y0, x0, xc, ty, pad_;
};

// Work item of the per-plane FFT passes: one (cell, component, x-plane).
struct PlaneItem {
  int cell, comp, x, pad_;
};

// Work item of the x-line FFT pass: a run of `len` consecutive (ky, kz) positions of one cell.
struct LineItem {
  int cell, pos0, len, pad_;
};

// Work item of the elementwise / reduction kernels: a chunk of one cell's voxels.
struct ChunkItem {
  int cell, pad_;
  int64_t begin, end;  // voxel range inside the cell
};

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }

// Sum of part[row][range(cell)] by one wavefront: lane-strided partial sums, then a fixed
// butterfly, so the result is deterministic; every lane returns the total.
__device__ __forceinline__ double wave_sum(const double* part, int nitems, int row, const int* range, int cell) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  const int end = range[2 * cell + 1];
  const double* pr = part + (int64_t)row * nitems;
  // four strides' loads in flight at once; the additions keep the order of the plain strided loop
  for (int w0 = range[2 * cell] + lane; w0 < end; w0 += 256) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = w0 + 64 * u < end ? pr[w0 + 64 * u] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (w0 + 64 * u < end) acc += v[u];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  return acc;
}

// Block-wide sum of one double per thread (blockDim.x a multiple of 64). Result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) t += red[i];
  }
  return t;
}

}  // namespace nlfh

// Error state of the C-ABI (thread-local message, status codes).
namespace nlfh {
void set_error(const std::string& msg);
}

#define NLFH_OK 0
#define NLFH_EINVAL 1
#define NLFH_EHIP 2
#define NLFH_EUNSUPPORTED 3
#define NLFH_ENOMEM 4

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) {                                                                \
      nlfh::set_error(std::string(#expr) + " failed: " + hipGetErrorString(_e) + " at " + \
                      __FILE__ + ":" + std::to_string(__LINE__));                          \
      return NLFH_EHIP;                                                                    \
    }                                                                                      \
  } while (0)
