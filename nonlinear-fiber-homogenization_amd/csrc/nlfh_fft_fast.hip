// Preconditioner passes for cubic cells of side N in {16, 32, 48, 64, 128} with register FFTs
// (nlfiber/linear.py:167-170). Same three passes and data layout as the generic kernels of
// nlfh_fft.hip; here every line transform is split over L threads: step 1 Dft<N1> in registers
// per thread, a padded LDS exchange, step 2 Dft<N2> in registers (N = N1 N2, LineSplit<N>).
//   A: per (cell, rhs, component, x): z real-to-complex (two rows per complex line), y complex
//      (the real DC and Nyquist columns packed into one complex line), store the half plane;
//      optionally fused with r -= alpha Ap, x += alpha p and ||r||^2.
//   B: per (cell, rhs, run of LPB (ky, kz) positions): x forward for the three components in
//      registers, z = Kinv r with the real symmetric symbol (component-planar copy), x inverse.
//   C: per (cell, rhs, component, x): y inverse, z complex-to-real, store z, optional r . z.
#include "nlfh_fft_codelets.h"
#include "nlfh_kernels.h"

namespace nlfh {

using fft::Dft;
using fft::LineSplit;

template <typename T, int N>
struct FastCfg {
  static constexpr int N1 = LineSplit<N>::N1, N2 = LineSplit<N>::N2;
  static constexpr int L = N1 > N2 ? N1 : N2;
  static constexpr int NH = N / 2 + 1;
  static constexpr int NP = N / 2;
  // threads of passes A, C; 128 takes 1024 so that one round of lines covers every row pair of a plane
  // (the staged FP32 passes need that; 139 KiB of LDS, one block per CU)
  static constexpr int BT = N == 128 ? 1024 : 256;
  // FP64 at 128: the plane alone takes 130 KiB, so a line's exchange between its two steps happens in the
  // line's own slots of the plane (a row pair or a column holds the line's N values; the L threads of a
  // line are in one wavefront, which reads all its slots before it writes them), with a skewed index in
  // place of padding; no separate exchange buffer
  static constexpr bool XIN = sizeof(T) == 8 && N == 128;
  static constexpr int LPR = BT / L;                                   // lines per round
  static constexpr int XS = N1 * (N2 + 1) > N2 * (N1 + 1) ? N1 * (N2 + 1) : N2 * (N1 + 1);
  // pass B positions per block: 256 threads (128 for FP64 at 128), one component's exchange buffer in LDS
  static constexpr int LPB = N == 128 ? (sizeof(T) == 8 ? 8 : 16) : 256 / L;
  static constexpr int BTB = L * LPB;                                          // threads of pass B
};

template <typename T>
__device__ __forceinline__ cplx<T> cmulw(cplx<T> v, cplx<T> w, bool conj_w) {
  const T wi = conj_w ? -w.im : w.im;
  return {v.re * w.re - v.im * wi, v.re * wi + v.im * w.re};
}

// Exchange slot of element (k1, t) of a line between its two steps (k1 < N1 outputs of step 1 of thread t < N2;
// step 2 thread t' reads (t', j2)): a padded row of a separate buffer, or, in place (XIN), position
// k1 N2 + ((t + k1) mod N2) of the line's own N slots of the plane, `stride` apart (1 for a row pair,
// NH for a column).
template <typename T, int N>
__device__ __forceinline__ int xslot(int k1, int t) {
  using C = FastCfg<T, N>;
  if constexpr (C::XIN) return k1 * C::N2 + ((t + k1) & (C::N2 - 1));
  else return k1 * (C::N2 + 1) + t;
}

// Pass A ---------------------------------------------------------------------------------------
template <typename T, int N, bool UPDATE>
__global__ void __launch_bounds__((FastCfg<T, N>::BT)) prec_a_fast(const PrecArgs<T> a, const cplx<T>* __restrict__ twg) {
  using C = FastCfg<T, N>;
  constexpr int N1 = C::N1, N2 = C::N2, NH = C::NH, NP = C::NP, L = C::L, LPR = C::LPR, BT = C::BT;
  extern __shared__ double smem_d[];
  __shared__ double red[16];  // one per wavefront (up to 1024 threads)
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_d);  // [N]: e^{-2 pi i m / N}
  cplx<T>* plane = tw + N;                            // [N][NH]
  cplx<T>* xch = plane + N * NH;                      // [LPR][XS]
  cplx<T>* dcn = xch + (C::XIN ? 0 : LPR * C::XS);     // [N]: packed DC / Nyquist column transform
  const int item = a.plane_base + blockIdx.x;
  const PlaneItem pi = a.planes[item];
  if (a.done != nullptr && a.done[pi.cell]) return;
  const CellDesc cd = a.cells[pi.cell];
  const int rhs = blockIdx.y;
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BT) tw[i] = twg[i];
  const int64_t nvox = cd.nvox;
  const int64_t fo = cd.off_field + ((int64_t)rhs * 3 + pi.comp) * nvox + (int64_t)pi.x * N * N;
  const T* __restrict__ R = a.r;
  T* __restrict__ RO = a.r_out;
  const T* __restrict__ AP = a.Ap;
  T* __restrict__ X = a.x;
  const T* __restrict__ P = a.p;
  T alpha = T(0);
  if (UPDATE) alpha = (T)a.alpha[pi.cell * a.m + rhs];
  const int l = tid / L, t = tid - (tid / L) * L;
  double acc = 0.0;
  __syncthreads();

  // z: complex transform of row pairs, unpacked to the two half spectra
  for (int p0 = 0; p0 < NP; p0 += LPR) {
    const int p = p0 + l;
    const bool act = p < NP;
    cplx<T>* xl = C::XIN ? plane + (2 * p) * NH : xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
      const int64_t r0 = fo + (int64_t)(2 * p) * N, r1 = r0 + N;
      T a0[N1], a1[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        a0[j1] = R[r0 + t + N2 * j1];
        a1[j1] = R[r1 + t + N2 * j1];
      }
      if (UPDATE) {
#pragma unroll
        for (int j1 = 0; j1 < N1; ++j1) {
          const int z = t + N2 * j1;
          a0[j1] -= alpha * AP[r0 + z];
          a1[j1] -= alpha * AP[r1 + z];
          RO[r0 + z] = a0[j1];
          RO[r1 + z] = a1[j1];
          X[r0 + z] += alpha * P[r0 + z];
          X[r1 + z] += alpha * P[r1 + z];
        }
      }
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        acc += (double)a0[j1] * (double)a0[j1] + (double)a1[j1] * (double)a1[j1];
        v[j1] = {a0[j1], a1[j1]};
      }
      Dft<T, N1, -1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[xslot<T, N>(k1, t)] = cmulw(v[k1], tw[t * k1], false);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[xslot<T, N>(t, j2)];
      Dft<T, N2, -1>::run(w);
      cplx<T>* zp = plane + (2 * p) * NH;  // the pair's two rows hold Z[0..N-1] for now
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) zp[t + N1 * k2] = w[k2];
    }
    __syncthreads();
    constexpr int ITEMS = LPR * NH;
    constexpr int IPT = (ITEMS + BT - 1) / BT;
    cplx<T> A[IPT], B[IPT];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int it = tid + i * BT;
      const int lp = it / NH, k = it - (it / NH) * NH;
      if (it < ITEMS && p0 + lp < NP) {
        const cplx<T>* zp = plane + (2 * (p0 + lp)) * NH;
        const cplx<T> Z = zp[k], M = zp[k == 0 ? 0 : N - k];
        A[i] = {T(0.5) * (Z.re + M.re), T(0.5) * (Z.im - M.im)};
        B[i] = {T(0.5) * (Z.im + M.im), T(-0.5) * (Z.re - M.re)};
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int it = tid + i * BT;
      const int lp = it / NH, k = it - (it / NH) * NH;
      if (it < ITEMS && p0 + lp < NP) {
        plane[(2 * (p0 + lp)) * NH + k] = A[i];
        plane[(2 * (p0 + lp) + 1) * NH + k] = B[i];
      }
    }
  }
  __syncthreads();

  // y: columns 1..NP-1 complex; columns 0 and NP (real) packed into line 0
  cplx<T>* so = a.spec + cd.off_spec + ((int64_t)rhs * 3 + pi.comp) * cd.nspec + (int64_t)pi.x * N * NH;
  for (int c0 = 0; c0 < NP; c0 += LPR) {
    const int col = c0 + l;
    const bool act = col < NP;
    cplx<T>* xl = C::XIN ? plane + col : xch + l * C::XS;
    constexpr int xs = C::XIN ? NH : 1;  // slot stride: a column of the plane in place
    if (act && t < N2) {
      cplx<T> v[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int y = t + N2 * j1;
        v[j1] = col == 0 ? cplx<T>{plane[y * NH].re, plane[y * NH + NP].re} : plane[y * NH + col];
      }
      Dft<T, N1, -1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[xslot<T, N>(k1, t) * xs] = cmulw(v[k1], tw[t * k1], false);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[xslot<T, N>(t, j2) * xs];
      Dft<T, N2, -1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int ky = t + N1 * k2;
        if (col == 0) dcn[ky] = w[k2];
        else so[(int64_t)ky * NH + col] = w[k2];
      }
    }
    __syncthreads();
  }
  // unpack the packed DC / Nyquist columns
  for (int ky = tid; ky < N; ky += BT) {
    const cplx<T> Z = dcn[ky], M = dcn[ky == 0 ? 0 : N - ky];
    so[(int64_t)ky * NH] = {T(0.5) * (Z.re + M.re), T(0.5) * (Z.im - M.im)};
    so[(int64_t)ky * NH + NP] = {T(0.5) * (Z.im + M.im), T(-0.5) * (Z.re - M.re)};
  }
  const double tsum = block_sum(acc, red);
  if (tid == 0 && a.part_a) a.part_a[(int64_t)rhs * a.nplanes + item] = tsum;
}

// Pass C ---------------------------------------------------------------------------------------
template <typename T, int N, bool DOT>
__global__ void __launch_bounds__((FastCfg<T, N>::BT)) prec_c_fast(const PrecArgs<T> a, const cplx<T>* __restrict__ twg) {
  using C = FastCfg<T, N>;
  constexpr int N1 = C::N1, N2 = C::N2, NH = C::NH, NP = C::NP, L = C::L, LPR = C::LPR, BT = C::BT;
  extern __shared__ double smem_d[];
  __shared__ double red[16];  // one per wavefront (up to 1024 threads)
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_d);
  cplx<T>* plane = tw + N;
  cplx<T>* xch = plane + N * NH;
  const int item = a.plane_base + blockIdx.x;
  const PlaneItem pi = a.planes[item];
  if (a.done != nullptr && a.done[pi.cell]) return;
  const CellDesc cd = a.cells[pi.cell];
  const int rhs = blockIdx.y;
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BT) tw[i] = twg[i];
  const cplx<T>* so = a.spec + cd.off_spec + ((int64_t)rhs * 3 + pi.comp) * cd.nspec + (int64_t)pi.x * N * NH;
  for (int i = tid; i < N * NH; i += BT) plane[i] = so[i];
  const int l = tid / L, t = tid - (tid / L) * L;
  __syncthreads();

  // y inverse; columns 0 and NP packed (their inverse transforms are real)
  for (int c0 = 0; c0 < NP; c0 += LPR) {
    const int col = c0 + l;
    const bool act = col < NP;
    cplx<T>* xl = C::XIN ? plane + col : xch + l * C::XS;
    constexpr int xs = C::XIN ? NH : 1;  // slot stride: a column of the plane in place
    if (act && t < N2) {
      cplx<T> v[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int y = t + N2 * j1;
        if (col == 0) {
          const cplx<T> A = plane[y * NH], B = plane[y * NH + NP];
          v[j1] = {A.re - B.im, A.im + B.re};
        } else {
          v[j1] = plane[y * NH + col];
        }
      }
      Dft<T, N1, 1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[xslot<T, N>(k1, t) * xs] = cmulw(v[k1], tw[t * k1], true);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[xslot<T, N>(t, j2) * xs];
      Dft<T, N2, 1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int y = t + N1 * k2;
        if (col == 0) {
          plane[y * NH] = {w[k2].re, T(0)};
          plane[y * NH + NP] = {w[k2].im, T(0)};
        } else {
          plane[y * NH + col] = w[k2];
        }
      }
    }
    __syncthreads();
  }

  // z inverse: row pairs, Z = A + i B with the Hermitian extension
  const int64_t nvox = cd.nvox;
  const int64_t fo = cd.off_field + ((int64_t)rhs * 3 + pi.comp) * nvox + (int64_t)pi.x * N * N;
  double acc = 0.0;
  for (int p0 = 0; p0 < NP; p0 += LPR) {
    const int p = p0 + l;
    const bool act = p < NP;
    cplx<T>* xl = C::XIN ? plane + (2 * p) * NH : xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
      const cplx<T>* ra = plane + (2 * p) * NH;
      const cplx<T>* rb = ra + NH;
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int k = t + N2 * j1;
        cplx<T> A, B;
        if (k < NH) {
          A = ra[k];
          B = rb[k];
        } else {
          A = ra[N - k];
          A.im = -A.im;
          B = rb[N - k];
          B.im = -B.im;
        }
        v[j1] = {A.re - B.im, A.im + B.re};
      }
      Dft<T, N1, 1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[xslot<T, N>(k1, t)] = cmulw(v[k1], tw[t * k1], true);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[xslot<T, N>(t, j2)];
      Dft<T, N2, 1>::run(w);
      const int64_t r0 = fo + (int64_t)(2 * p) * N, r1 = r0 + N;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int z = t + N1 * k2;
        a.z[r0 + z] = w[k2].re;
        a.z[r1 + z] = w[k2].im;
        if (DOT) acc += (double)a.r[r0 + z] * (double)w[k2].re + (double)a.r[r1 + z] * (double)w[k2].im;
      }
    }
    __syncthreads();
  }
  if (DOT) {
    const double tsum = block_sum(acc, red);
    if (tid == 0) a.part_c[(int64_t)rhs * a.nplanes + item] = tsum;
  }
}

// Passes A and C with the x-plane staged through LDS (sizes whose lines all fit one round,
// LPR >= N / 2): every global access is a 16-byte vector over consecutive z, so a wavefront reads or
// writes 1 KiB contiguous per instruction. The real plane [N][RS] (RS = N + pad, keeps the row-pair
// reads of the z transform free of bank conflicts) aliases the exchange buffer: it is read completely
// before the exchange is written (pass A) and written only after the exchange is read (pass C).
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
  __device__ static float get(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
  __device__ static double get(const double2& v, int k) { return k == 0 ? v.x : v.y; }
};

// IP, the in-place form: each line's exchange in its own slots of the plane (a row pair or a column, as the FP64
// 128-point kernels do) and the real plane staged in the spectrum's own slots (row y in row y's 2 NH values), so a
// block needs the plane alone. FP32 at 128 (with B2: 512-thread blocks over two rounds of lines): 69 KiB, two blocks
// per CU, where the separate exchange buffer (139 KiB) let one 1024-thread block hold the CU and its load and
// transform phases could not overlap another block's. FP64 at 64: 36 KiB, four blocks per CU instead of two.
template <typename T, int N>
struct StageCfg {
  static constexpr bool IP = (sizeof(T) == 4 && N == 128) || (sizeof(T) == 8 && N == 64);
  // B2: two rounds of lines in 512-thread blocks, operands read where they are consumed (FP32 at 128)
  static constexpr bool B2 = sizeof(T) == 4 && N == 128;
  static constexpr int BT = B2 ? 512 : FastCfg<T, N>::BT;
  static constexpr int LPR = BT / FastCfg<T, N>::L;
  static constexpr int RS = IP ? 2 * FastCfg<T, N>::NH : N + (sizeof(T) == 4 ? 4 : 2);
  static constexpr bool fits = IP || (!FastCfg<T, N>::XIN && FastCfg<T, N>::LPR >= N / 2 &&
                               (size_t)N * RS * sizeof(T) <= (size_t)FastCfg<T, N>::LPR * FastCfg<T, N>::XS * sizeof(cplx<T>));
  // waves per SIMD the compiler may plan for: the in-place form's LDS admits two 8-wave blocks per CU (4 per
  // SIMD), so it gets their 128 registers; the others keep the default range
  static constexpr int WMIN = B2 ? 4 : 1, WMAX = B2 ? 4 : 8;
  static constexpr size_t lds_bytes =
      IP ? ((size_t)N + (size_t)N * FastCfg<T, N>::NH + N) * sizeof(cplx<T>)
         : ((size_t)N + (size_t)N * FastCfg<T, N>::NH + (size_t)FastCfg<T, N>::LPR * FastCfg<T, N>::XS + N) * sizeof(cplx<T>);
};

// exchange slot of element (k1, t) of a line inside the line's own N slots (in-place form of xslot)
template <int N2>
__device__ __forceinline__ int xslot_ip(int k1, int t) {
  return k1 * N2 + ((t + k1) & (N2 - 1));
}

// UPD: 0 none; 1 r -= alpha Ap; 2 also x += alpha p; 4 spectral residual (FP32 inner iteration): the
// plane of Ap is transformed and the residual's (z, y) spectrum takes the update in place,
// r-hat -= alpha Ap-hat, with ||r||^2 from that spectrum by Parseval (weight 1 at kz = 0 and N/2, 2
// elsewhere, over N^2), so r itself is never read or written; alpha = (r . z) / (p . Ap) is formed by every
// block from the matvec's partials with the fixed-order sum and guard of cg_alpha_kernel (bitwise its
// value), and the block of plane 0, component 0 stores it for pass C. Without UPD 4, the plane's spectrum
// goes to a.rspec when that is set (the start of such an iteration), else to a.spec.
// PPB (spectral update only): x-planes per block. With 2, a block transforms two consecutive x-planes of one
// (cell, component) in turn and issues the loads of the second (Ap and the residual spectrum) as soon as the first
// is staged in LDS, so they are in flight during the first plane's transforms (the passes are LDS-bound to four
// blocks per CU, which otherwise leaves their loads to the blocks' own load phases).
template <typename T, int N, int UPD, int PPB = 1>
__global__ void __launch_bounds__((StageCfg<T, N>::BT)) __attribute__((amdgpu_waves_per_eu(StageCfg<T, N>::WMIN, StageCfg<T, N>::WMAX))) prec_a_staged(const PrecArgs<T> a, const cplx<T>* __restrict__ twg) {
  constexpr bool UPDATE = UPD == 1 || UPD == 2;
  constexpr bool SPEC = UPD == 4;
  static_assert(PPB == 1 || (PPB == 2 && SPEC), "two planes per block: spectral update only");
  using C = FastCfg<T, N>;
  using SC = StageCfg<T, N>;
  constexpr int N1 = C::N1, N2 = C::N2, NH = C::NH, NP = C::NP, L = C::L, BT = SC::BT, LPR = SC::LPR;
  constexpr bool IP = SC::IP, B2 = SC::B2;
  static_assert(!IP || PPB == 1, "one plane per block in the in-place form");
  constexpr int RS = StageCfg<T, N>::RS;
  static_assert(StageCfg<T, N>::fits, "staged pass A needs one round of lines and room for the plane");
  using VT = Vec16<T>;
  constexpr int VN = VT::n;
  extern __shared__ double smem_d[];
  __shared__ double red[16];  // one per wavefront (up to 1024 threads)
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_d);
  cplx<T>* plane = tw + N;
  cplx<T>* xch = plane + N * NH;  // not used in the in-place form
  cplx<T>* dcn = IP ? xch : xch + C::LPR * C::XS;
  T* rp = IP ? reinterpret_cast<T*>(plane) : reinterpret_cast<T*>(xch);
  const int item0 = a.plane_base + (int)blockIdx.x * PPB;
  if (PPB > 1 && (int)blockIdx.x * PPB >= a.plane_count) return;
  const PlaneItem pi0 = a.planes[item0];
  if (a.done != nullptr && a.done[pi0.cell]) return;
  const CellDesc cd = a.cells[pi0.cell];
  const int rhs = blockIdx.y;
  const int tid = threadIdx.x;
  // twiddles: loaded into a register here, stored to LDS with the plane (one wait for all loads)
  static_assert(BT >= N, "one twiddle per thread");
  cplx<T> twv{};
  if (tid < N) twv = twg[tid];
  using V = typename VT::type;
  constexpr int NV = N * N / VN;
  // without the real-space update (UPD 0 and 4) a thread's whole share of the plane is loaded into
  // registers first, so that its loads are in flight together (a plain loop waits for each load before
  // the LDS store that follows it); they are issued ahead of the other loads, which are consumed later
  constexpr bool BATCH = !UPDATE;
  constexpr int LPT = BATCH ? (NV + BT - 1) / BT : 1;
  // SPEC: the plane's residual spectrum is loaded with it, into registers, so that its latency passes
  // during the transforms (it is consumed by the final update)
  constexpr int CPV = 16 / sizeof(cplx<T>);  // complex values per vector (2 float, 1 double)
  constexpr int NVS = N * NH / CPV;
  constexpr int RPT = SPEC ? (NVS + BT - 1) / BT : 1;
  V lbuf[PPB][LPT];
  V rpre[PPB][B2 ? 1 : RPT];
  auto field_off = [&](const PlaneItem& q) {
    return cd.off_field + ((int64_t)rhs * 3 + q.comp) * cd.nvox + (int64_t)q.x * N * N;
  };
  auto spec_ptr = [&](const PlaneItem& q) {
    return reinterpret_cast<V*>((a.rspec != nullptr ? a.rspec : a.spec) + cd.off_spec +
                                ((int64_t)rhs * 3 + q.comp) * cd.nspec + (int64_t)q.x * N * NH);
  };
  auto issue = [&](int q, const PlaneItem& pq) {
    if constexpr (BATCH) {
      const V* Rv = reinterpret_cast<const V*>((SPEC ? a.Ap : a.r) + field_off(pq));
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int i = tid + j * BT;
        if (i < NV) lbuf[q][j] = Rv[i];
      }
    }
    if constexpr (SPEC && !B2) {
      const V* rs = spec_ptr(pq);
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int i = tid + j * BT;
        if (i < NVS) rpre[q][j] = rs[i];
      }
    }
  };
  issue(0, pi0);
  T alpha = T(0);
  __shared__ double alpha_sh;
  if constexpr (UPDATE) {
    // alpha from the matvec's p . Ap partials (a.part_mv set: no alpha kernel), with the fixed-order sum and guard
    // of cg_alpha_kernel, so bitwise its value; the block of plane 0, component 0 stores it (and the one it
    // replaces); else the value the alpha kernel stored
    if (a.part_mv != nullptr) {
      if (tid < 64) {
        const int k = pi0.cell * a.m + rhs;
        const double pAp = wave_sum(a.part_mv, a.nmv, rhs, a.mv_range, pi0.cell);
        const double al = a.rz[k] / (pAp == 0.0 ? 1.0 : pAp);
        if (tid == 0) {
          alpha_sh = al;
          if (pi0.x == 0 && pi0.comp == 0) {
            if (a.alpha_prev_out != nullptr) a.alpha_prev_out[k] = a.alpha_out[k];
            a.alpha_out[k] = al;
          }
        }
      }
      __syncthreads();
      alpha = (T)alpha_sh;
    } else {
      alpha = (T)a.alpha[pi0.cell * a.m + rhs];
    }
  }
  if (SPEC && tid < 64) {
    const int k = pi0.cell * a.m + rhs;
    const double pAp = wave_sum(a.part_mv, a.nmv, rhs, a.mv_range, pi0.cell);
    const double al = a.rz[k] / (pAp == 0.0 ? 1.0 : pAp);
    if (tid == 0) {
      alpha_sh = al;
      if (pi0.x == 0 && pi0.comp == 0) {  // no other block of this pass reads the stored values
        a.alpha_prev_out[k] = a.alpha_out[k];
        a.alpha_out[k] = al;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPB; ++q) {
  const int item = item0 + q;
  if (PPB > 1 && q > 0 && (int)blockIdx.x * PPB + q >= a.plane_count) break;
  const PlaneItem pi = q == 0 ? pi0 : a.planes[item];
  const int64_t fo = field_off(pi);
  const V* Rv = reinterpret_cast<const V*>((SPEC ? a.Ap : a.r) + fo);
  V* const rsp = spec_ptr(pi);
  double acc = 0.0;
  // load (and update) the plane with 16-byte vectors into rp
  {
    V* ROv = reinterpret_cast<V*>(a.r_out + fo);
    const V* APv = reinterpret_cast<const V*>(a.Ap + fo);
    V* Xv = reinterpret_cast<V*>(a.x + fo);
    const V* Pv = reinterpret_cast<const V*>(a.p + fo);
    if (q == 0 && tid < N) tw[tid] = twv;
    // vector i of the plane into the staged real plane (and its norm)
    auto put = [&](int i, const V& rv) {
      const int e = i * VN;
      const int y = e / N, z = e - (e / N) * N;
      T* dst = rp + y * RS + z;
#pragma unroll
      for (int k = 0; k < VN; ++k) {
        const T v = VT::get(rv, k);
        if (!SPEC) acc += (double)v * (double)v;
        dst[k] = v;
      }
    };
    if constexpr (BATCH) {
#pragma unroll
      for (int j = 0; j < LPT; ++j)
        if (tid + j * BT < NV) put(tid + j * BT, lbuf[q][j]);
      // the next plane's loads go out now, to be in flight during this plane's transforms
      if (q + 1 < PPB && (int)blockIdx.x * PPB + q + 1 < a.plane_count) issue(q + 1, a.planes[item + 1]);
    } else {
      for (int i = tid; i < NV; i += BT) {
        V rv = Rv[i];
        const V ap = APv[i];
        if constexpr (VN == 4) {
          rv.x -= alpha * ap.x; rv.y -= alpha * ap.y; rv.z -= alpha * ap.z; rv.w -= alpha * ap.w;
        } else {
          rv.x -= alpha * ap.x; rv.y -= alpha * ap.y;
        }
        ROv[i] = rv;
        if (UPD >= 2) {
          const V pv = Pv[i];
          V xv = Xv[i];
          if constexpr (VN == 4) {
            xv.x += alpha * pv.x; xv.y += alpha * pv.y; xv.z += alpha * pv.z; xv.w += alpha * pv.w;
          } else {
            xv.x += alpha * pv.x; xv.y += alpha * pv.y;
          }
          Xv[i] = xv;
        }
        put(i, rv);
      }
    }
  }
  const int l = tid / L, t = tid - (tid / L) * L;
  __syncthreads();

  // z: complex transform of row pairs (line l = pair p; two rounds in the in-place form, where a pair's exchange
  // and its Z stay in the pair's own rows, which only the pair's line (one wavefront) reads and writes)
  {
  if constexpr (IP) {
  for (int p0 = 0; p0 < (IP ? NP : 1); p0 += LPR) {  // one pass over all pairs when not in place
    const int p = IP ? p0 + l : l;
    const bool act = p < NP && t < N2;
    cplx<T> v[N1];
    if (act) {
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) v[j1] = {rp[(2 * p) * RS + t + N2 * j1], rp[(2 * p + 1) * RS + t + N2 * j1]};
    }
    __syncthreads();  // rp is read completely before the exchange (same memory) is written
    cplx<T>* xl = IP ? plane + (2 * p) * NH : xch + l * C::XS;
    if (act) {
      Dft<T, N1, -1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1)
        xl[IP ? xslot_ip<N2>(k1, t) : k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], false);
    }
    __syncthreads();
    if (p < NP && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[IP ? xslot_ip<N2>(t, j2) : t * (N2 + 1) + j2];
      Dft<T, N2, -1>::run(w);
      cplx<T>* zp = plane + (2 * p) * NH;  // the pair's two rows hold Z[0..N-1] for now
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) zp[t + N1 * k2] = w[k2];
    }
    __syncthreads();
  }
  } else {
    const int p = l;
    const bool act = p < NP && t < N2;
    cplx<T> v[N1];
    if (act) {
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) v[j1] = {rp[(2 * p) * RS + t + N2 * j1], rp[(2 * p + 1) * RS + t + N2 * j1]};
    }
    __syncthreads();  // rp is read completely before the exchange (same memory) is written
    cplx<T>* xl = xch + l * C::XS;
    if (act) {
      Dft<T, N1, -1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], false);
    }
    __syncthreads();
    if (p < NP && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[t * (N2 + 1) + j2];
      Dft<T, N2, -1>::run(w);
      cplx<T>* zp = plane + (2 * p) * NH;  // the pair's two rows hold Z[0..N-1] for now
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) zp[t + N1 * k2] = w[k2];
    }
    __syncthreads();
  }
    constexpr int ITEMS = NP * NH;
    constexpr int IPT = (ITEMS + BT - 1) / BT;
    cplx<T> A[IPT], B[IPT];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int it = tid + i * BT;
      const int lp = it / NH, k = it - (it / NH) * NH;
      if (it < ITEMS) {
        const cplx<T>* zq = plane + (2 * lp) * NH;
        const cplx<T> Z = zq[k], M = zq[k == 0 ? 0 : N - k];
        A[i] = {T(0.5) * (Z.re + M.re), T(0.5) * (Z.im - M.im)};
        B[i] = {T(0.5) * (Z.im + M.im), T(-0.5) * (Z.re - M.re)};
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int it = tid + i * BT;
      const int lp = it / NH, k = it - (it / NH) * NH;
      if (it < ITEMS) {
        plane[(2 * lp) * NH + k] = A[i];
        plane[(2 * lp + 1) * NH + k] = B[i];
      }
    }
  }
  __syncthreads();

  // y: columns 1..NP-1 complex, columns 0 and NP (real) packed into line 0; results in place (the in-place form
  // exchanges inside the column, stride NH)
  if constexpr (IP) {
  for (int c0 = 0; c0 < (IP ? NP : 1); c0 += LPR) {  // one pass over all columns when not in place
    const int col = IP ? c0 + l : l;
    const bool act = col < NP;
    cplx<T>* xl = IP ? plane + col : xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int y = t + N2 * j1;
        v[j1] = col == 0 ? cplx<T>{plane[y * NH].re, plane[y * NH + NP].re} : plane[y * NH + col];
      }
      Dft<T, N1, -1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1)
        xl[IP ? xslot_ip<N2>(k1, t) * NH : k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], false);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[IP ? xslot_ip<N2>(t, j2) * NH : t * (N2 + 1) + j2];
      Dft<T, N2, -1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int ky = t + N1 * k2;
        if (col == 0) dcn[ky] = w[k2];
        else plane[ky * NH + col] = w[k2];
      }
    }
    __syncthreads();
  }
  } else {
  {
    const int col = l;
    const bool act = col < NP;
    cplx<T>* xl = xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int y = t + N2 * j1;
        v[j1] = col == 0 ? cplx<T>{plane[y * NH].re, plane[y * NH + NP].re} : plane[y * NH + col];
      }
      Dft<T, N1, -1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], false);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[t * (N2 + 1) + j2];
      Dft<T, N2, -1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int ky = t + N1 * k2;
        if (col == 0) dcn[ky] = w[k2];
        else plane[ky * NH + col] = w[k2];
      }
    }
    __syncthreads();
  }
  }
  // unpack the packed DC / Nyquist columns
  for (int ky = tid; ky < N; ky += BT) {
    const cplx<T> Z = dcn[ky], M = dcn[ky == 0 ? 0 : N - ky];
    plane[ky * NH] = {T(0.5) * (Z.re + M.re), T(0.5) * (Z.im - M.im)};
    plane[ky * NH + NP] = {T(0.5) * (Z.im + M.im), T(-0.5) * (Z.re - M.re)};
  }
  __syncthreads();
  // the half plane [ky][kz] is contiguous in the spectrum: 16-byte vector copy
  {
    const V* src = reinterpret_cast<const V*>(plane);
    V* dst = rsp;
    if constexpr (SPEC) {
      alpha = (T)alpha_sh;  // written before the barriers above
      double sacc = 0.0;
      if constexpr (!B2) {
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = tid + j * BT;
          if (i >= NVS) break;
          V rh = rpre[q][j];
          const V ap = src[i];
          if constexpr (VN == 4) {
            rh.x -= alpha * ap.x; rh.y -= alpha * ap.y; rh.z -= alpha * ap.z; rh.w -= alpha * ap.w;
          } else {
            rh.x -= alpha * ap.x; rh.y -= alpha * ap.y;
          }
          dst[i] = rh;
#pragma unroll
          for (int c = 0; c < CPV; ++c) {
            const int kz = (i * CPV + c) % NH;
            const double re = (double)VT::get(rh, 2 * c), im = (double)VT::get(rh, 2 * c + 1);
            sacc += (kz == 0 || 2 * kz == N ? 1.0 : 2.0) * (re * re + im * im);
          }
        }
      } else {
        // in-place form: the residual spectrum is read here (not prefetched; the other block on the CU covers the
        // latency), its loads issued together
        const V* rs = spec_ptr(pi);
        V rh[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = tid + j * BT;
          if (i < NVS) rh[j] = rs[i];
        }
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = tid + j * BT;
          if (i < NVS) {
            const V ap = src[i];
            rh[j].x -= alpha * ap.x; rh[j].y -= alpha * ap.y; rh[j].z -= alpha * ap.z; rh[j].w -= alpha * ap.w;
            dst[i] = rh[j];
#pragma unroll
            for (int c = 0; c < CPV; ++c) {
              const int kz = (i * CPV + c) % NH;
              const double re = (double)VT::get(rh[j], 2 * c), im = (double)VT::get(rh[j], 2 * c + 1);
              sacc += (kz == 0 || 2 * kz == N ? 1.0 : 2.0) * (re * re + im * im);
            }
          }
        }
      }
      acc = sacc * (1.0 / ((double)N * (double)N));
    } else {
      for (int i = tid; i < NVS; i += BT) dst[i] = src[i];
    }
  }
  const double tsum = block_sum(acc, red);
  if (tid == 0 && a.part_a) a.part_a[(int64_t)rhs * a.nplanes + item] = tsum;
  if (PPB > 1) __syncthreads();  // the next plane reuses the LDS
  }
}

// DIR (FP32 inner iteration): instead of z, the new search direction p_{k+1} = z + beta_k p_k goes to
// a.z (p_k = a.p, the direction of this iteration's matvec). beta_k = (r . z)_{k+1} / (r . z)_k is formed
// by every block from pass B's partials with the same fixed-order sum (wave_sum) and guard as
// cg_beta_kernel, so it is bitwise the value cg_beta stores. The matvec then reads p alone: z is neither
// written nor read again. DIR 1 (even k) leaves the iterate alone; DIR 2 (odd k) gives it the steps of
// both iterations, x += alpha_{k-1} p_{k-1} + alpha_k p_k, where p_{k-1} is the old content of a.z, read
// by the same thread before it writes p_{k+1} there (x is read and written every second iteration).
template <typename T, int N, bool DOT, int DIR = 0>
__global__ void __launch_bounds__((StageCfg<T, N>::BT)) __attribute__((amdgpu_waves_per_eu(StageCfg<T, N>::WMIN, StageCfg<T, N>::WMAX))) prec_c_staged(const PrecArgs<T> a, const cplx<T>* __restrict__ twg) {
  static_assert(!(DOT && DIR), "the direction pass takes r . z from pass B");
  using C = FastCfg<T, N>;
  using SC = StageCfg<T, N>;
  constexpr int N1 = C::N1, N2 = C::N2, NH = C::NH, NP = C::NP, L = C::L, BT = SC::BT, LPR = SC::LPR;
  constexpr bool IP = SC::IP, B2 = SC::B2;
  constexpr int RS = StageCfg<T, N>::RS;
  static_assert(StageCfg<T, N>::fits, "staged pass C needs one round of lines and room for the plane");
  using VT = Vec16<T>;
  using V = typename VT::type;
  constexpr int VN = VT::n;
  extern __shared__ double smem_d[];
  __shared__ double red[16];  // one per wavefront (up to 1024 threads)
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_d);
  cplx<T>* plane = tw + N;
  cplx<T>* xch = plane + N * NH;  // not used in the in-place form
  T* rp = IP ? reinterpret_cast<T*>(plane) : reinterpret_cast<T*>(xch);
  const int item = a.plane_base + blockIdx.x;
  const PlaneItem pi = a.planes[item];
  if (a.done != nullptr && a.done[pi.cell]) return;
  const CellDesc cd = a.cells[pi.cell];
  const int rhs = blockIdx.y;
  const int tid = threadIdx.x;
  // twiddles: loaded into a register here, stored to LDS with the plane (one wait for all loads)
  static_assert(BT >= N, "one twiddle per thread");
  cplx<T> twv{};
  if (tid < N) twv = twg[tid];
  // a thread's share of the spectrum plane is loaded into registers first, ahead of the
  // other loads (its loads in flight together; a plain loop waits for each load before the LDS store that
  // follows it); at 128 the odd-iteration direction variant has no registers to spare and keeps the plain loop
  constexpr int CPV = 16 / sizeof(cplx<T>);  // complex values per vector (2 float, 1 double)
  constexpr int NVS = N * NH / CPV;
  constexpr bool SBATCH = N <= 64 || DIR != 2;
  constexpr int SPT = SBATCH ? (NVS + BT - 1) / BT : 1;
  const V* const ssrc = reinterpret_cast<const V*>(a.spec + cd.off_spec + ((int64_t)rhs * 3 + pi.comp) * cd.nspec +
                                                   (int64_t)pi.x * N * NH);
  V sbuf[SPT];
  if constexpr (SBATCH) {
#pragma unroll
    for (int j = 0; j < SPT; ++j)
      if (tid + j * BT < NVS) sbuf[j] = ssrc[tid + j * BT];
  }
  __shared__ double beta_sh;
  if (DIR && tid < 64) {
    const double rz1 = wave_sum(a.part_bdot, a.nlines, rhs, a.line_range, pi.cell);
    const double rz = a.rz[pi.cell * a.m + rhs];
    if (tid == 0) beta_sh = rz1 / (rz == 0.0 ? 1.0 : rz);
  }
  // DIR: the plane's p (and with DIR 2 x and p_{k-1}) are loaded now, into registers, so that their
  // latency passes during the transforms (they are consumed by the final store)
  const int64_t fo = cd.off_field + ((int64_t)rhs * 3 + pi.comp) * cd.nvox + (int64_t)pi.x * N * N;
  constexpr int NV = N * N / VN;
  constexpr int PPT = DIR ? (NV + BT - 1) / BT : 1;
  // DIR 2 in the in-place form (twice the values per thread) reads x and p_{k-1} at the store instead
  constexpr bool XQPRE = DIR == 2 && !B2;
  constexpr int XPT = XQPRE ? PPT : 1;
  V ppre[PPT], xpre[XPT], qpre[XPT];
  if constexpr (DIR != 0) {
    const V* Pv = reinterpret_cast<const V*>(a.p + fo);
    const V* Xv = reinterpret_cast<const V*>(a.x + fo);
    const V* Qv = reinterpret_cast<const V*>(a.z + fo);
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int i = tid + j * BT;
      if (i < NV) {
        ppre[j] = Pv[i];
        if constexpr (XQPRE) {
          xpre[j] = Xv[i];
          qpre[j] = Qv[i];
        }
      }
    }
  }
  {
    V* dst = reinterpret_cast<V*>(plane);
    if (tid < N) tw[tid] = twv;
    if constexpr (SBATCH) {
#pragma unroll
      for (int j = 0; j < SPT; ++j)
        if (tid + j * BT < NVS) {
          V* d = dst + tid + j * BT;
          if constexpr (CPV == 2) {
            d->x = sbuf[j].x; d->y = sbuf[j].y; d->z = sbuf[j].z; d->w = sbuf[j].w;
          } else {
            d->x = sbuf[j].x; d->y = sbuf[j].y;
          }
        }
    } else {
      for (int i = tid; i < NVS; i += BT) dst[i] = ssrc[i];
    }
  }
  const int l = tid / L, t = tid - (tid / L) * L;
  __syncthreads();

  // y inverse; columns 0 and NP packed (their inverse transforms are real); in place (two rounds of columns in
  // the in-place form, whose exchange stays in the column)
  if constexpr (IP) {
  for (int c0 = 0; c0 < (IP ? NP : 1); c0 += LPR) {  // one pass over all columns when not in place
    const int col = IP ? c0 + l : l;
    const bool act = col < NP;
    cplx<T>* xl = IP ? plane + col : xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int y = t + N2 * j1;
        if (col == 0) {
          const cplx<T> A = plane[y * NH], B = plane[y * NH + NP];
          v[j1] = {A.re - B.im, A.im + B.re};
        } else {
          v[j1] = plane[y * NH + col];
        }
      }
      Dft<T, N1, 1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1)
        xl[IP ? xslot_ip<N2>(k1, t) * NH : k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], true);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[IP ? xslot_ip<N2>(t, j2) * NH : t * (N2 + 1) + j2];
      Dft<T, N2, 1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int y = t + N1 * k2;
        if (col == 0) {
          plane[y * NH] = {w[k2].re, T(0)};
          plane[y * NH + NP] = {w[k2].im, T(0)};
        } else {
          plane[y * NH + col] = w[k2];
        }
      }
    }
    __syncthreads();
  }

  } else {
  {
    const int col = l;
    const bool act = col < NP;
    cplx<T>* xl = xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int y = t + N2 * j1;
        if (col == 0) {
          const cplx<T> A = plane[y * NH], B = plane[y * NH + NP];
          v[j1] = {A.re - B.im, A.im + B.re};
        } else {
          v[j1] = plane[y * NH + col];
        }
      }
      Dft<T, N1, 1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], true);
    }
    __syncthreads();
    if (act && t < N1) {
      cplx<T> w[N2];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[t * (N2 + 1) + j2];
      Dft<T, N2, 1>::run(w);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int y = t + N1 * k2;
        if (col == 0) {
          plane[y * NH] = {w[k2].re, T(0)};
          plane[y * NH + NP] = {w[k2].im, T(0)};
        } else {
          plane[y * NH + col] = w[k2];
        }
      }
    }
    __syncthreads();
  }

  }
  // z inverse: row pairs, Z = A + i B with the Hermitian extension; outputs into rp (in the in-place form the
  // pair's exchange and its two real output rows stay in the pair's own rows)
  if constexpr (IP) {
  for (int p0 = 0; p0 < (IP ? NP : 1); p0 += LPR) {  // one pass over all pairs when not in place
    const int p = IP ? p0 + l : l;
    const bool act = p < NP;
    cplx<T>* xl = IP ? plane + (2 * p) * NH : xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
      const cplx<T>* ra = plane + (2 * p) * NH;
      const cplx<T>* rb = ra + NH;
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int k = t + N2 * j1;
        cplx<T> A, B;
        if (k < NH) {
          A = ra[k];
          B = rb[k];
        } else {
          A = ra[N - k];
          A.im = -A.im;
          B = rb[N - k];
          B.im = -B.im;
        }
        v[j1] = {A.re - B.im, A.im + B.re};
      }
      Dft<T, N1, 1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1)
        xl[IP ? xslot_ip<N2>(k1, t) : k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], true);
    }
    __syncthreads();
    cplx<T> w[N2];
    if (act && t < N1) {
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[IP ? xslot_ip<N2>(t, j2) : t * (N2 + 1) + j2];
      Dft<T, N2, 1>::run(w);
    }
    __syncthreads();  // the exchange is read completely before rp (same memory) is written
    if (act && t < N1) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int z = t + N1 * k2;
        rp[(2 * p) * RS + z] = w[k2].re;
        rp[(2 * p + 1) * RS + z] = w[k2].im;
      }
    }
    __syncthreads();
  }
  } else {
  {
    const int p = l;
    const bool act = p < NP;
    cplx<T>* xl = xch + l * C::XS;
    if (act && t < N2) {
      cplx<T> v[N1];
      const cplx<T>* ra = plane + (2 * p) * NH;
      const cplx<T>* rb = ra + NH;
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int k = t + N2 * j1;
        cplx<T> A, B;
        if (k < NH) {
          A = ra[k];
          B = rb[k];
        } else {
          A = ra[N - k];
          A.im = -A.im;
          B = rb[N - k];
          B.im = -B.im;
        }
        v[j1] = {A.re - B.im, A.im + B.re};
      }
      Dft<T, N1, 1>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xl[k1 * (N2 + 1) + t] = cmulw(v[k1], tw[t * k1], true);
    }
    __syncthreads();
    cplx<T> w[N2];
    if (act && t < N1) {
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) w[j2] = xl[t * (N2 + 1) + j2];
      Dft<T, N2, 1>::run(w);
    }
    __syncthreads();  // the exchange is read completely before rp (same memory) is written
    if (act && t < N1) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int z = t + N1 * k2;
        rp[(2 * p) * RS + z] = w[k2].re;
        rp[(2 * p + 1) * RS + z] = w[k2].im;
      }
    }
    __syncthreads();
  }
  }
  // store z (and r . z) with 16-byte vectors; DIR: p_new = z + beta p and x += alpha p instead
  double acc = 0.0;
  {
    V* Zv = reinterpret_cast<V*>(a.z + fo);
    const V* Rv = reinterpret_cast<const V*>(a.r + fo);
    V* Xv = reinterpret_cast<V*>(DIR ? a.x + fo : nullptr);
    T beta = T(0), alpha = T(0), alpha_p = T(0);
    if (DIR) beta = (T)beta_sh;  // written before the barriers above
    if (DIR == 2) {
      alpha = (T)a.alpha[pi.cell * a.m + rhs];
      alpha_p = (T)a.alpha_prev[pi.cell * a.m + rhs];
    }
    // vector i of the plane; j: its slot among the prefetched values (DIR)
    auto store = [&](int i, int j) {
      const int e = i * VN;
      const int y = e / N, z = e - (e / N) * N;
      const T* srcp = rp + y * RS + z;
      V zv;
      if constexpr (VN == 4) zv = {srcp[0], srcp[1], srcp[2], srcp[3]};
      else zv = {srcp[0], srcp[1]};
      if constexpr (DIR != 0) {
        const V pv = ppre[j];
        if constexpr (VN == 4) {
          zv.x += beta * pv.x; zv.y += beta * pv.y; zv.z += beta * pv.z; zv.w += beta * pv.w;
        } else {
          zv.x += beta * pv.x; zv.y += beta * pv.y;
        }
        if constexpr (DIR == 2) {
          const V qv = XQPRE ? qpre[XQPRE ? j : 0] : reinterpret_cast<const V*>(a.z + fo)[i];
          V xv = XQPRE ? xpre[XQPRE ? j : 0] : Xv[i];
          if constexpr (VN == 4) {
            xv.x += alpha_p * qv.x; xv.y += alpha_p * qv.y; xv.z += alpha_p * qv.z; xv.w += alpha_p * qv.w;
            xv.x += alpha * pv.x; xv.y += alpha * pv.y; xv.z += alpha * pv.z; xv.w += alpha * pv.w;
          } else {
            xv.x += alpha_p * qv.x; xv.y += alpha_p * qv.y;
            xv.x += alpha * pv.x; xv.y += alpha * pv.y;
          }
          Xv[i] = xv;
        }
      }
      Zv[i] = zv;
      if (DOT) {
        const V rv = Rv[i];
#pragma unroll
        for (int k = 0; k < VN; ++k) acc += (double)VT::get(rv, k) * (double)VT::get(zv, k);
      }
    };
    if constexpr (DIR != 0) {
#pragma unroll
      for (int j = 0; j < PPT; ++j)
        if (tid + j * BT < NV) store(tid + j * BT, j);
    } else {
      for (int i = tid; i < NV; i += BT) store(i, 0);
    }
  }
  if (DOT) {
    const double tsum = block_sum(acc, red);
    if (tid == 0) a.part_c[(int64_t)rhs * a.nplanes + item] = tsum;
  }
}

// Pass B ---------------------------------------------------------------------------------------
// DOTB: also the partial r . z of the line item from the spectra (Parseval over the half spectrum:
// weight 1 at kz = 0 and kz = N/2, 2 elsewhere; z-hat = Kinv r-hat / N is what the inverse transform
// turns into z), so pass C need not read r.
// LPBX: (ky, kz) positions per block when not the default C::LPB (a launch shape the autotuner picks; the plan's
// line items then hold LPBX positions each), with L * LPBX threads.
template <typename T, int N, int LPBX>
constexpr int pass_b_threads() {
  return LPBX > 0 ? FastCfg<T, N>::L * LPBX : FastCfg<T, N>::BTB;
}

template <typename T, int N, bool DOTB, int LPBX = 0>
__global__ void __launch_bounds__((pass_b_threads<T, N, LPBX>())) prec_b_fast(const PrecArgs<T> a,
                                                                             const cplx<T>* __restrict__ twg,
                                                                             const T* __restrict__ symp) {
  __shared__ double red[16];  // one per wavefront (up to 1024 threads)
  double dacc = 0.0;
  using C = FastCfg<T, N>;
  constexpr int N1 = C::N1, N2 = C::N2, NH = C::NH;
  constexpr int LPB = LPBX > 0 ? LPBX : C::LPB, BT = pass_b_threads<T, N, LPBX>();
  extern __shared__ double smem_d[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_d);
  cplx<T>* xch = tw + N;  // [XS][LPB], one component at a time
  const LineItem li = a.lines[a.line_base + blockIdx.x];
  if (a.done != nullptr && a.done[li.cell]) return;
  const CellDesc cd = a.cells[li.cell];
  const int rhs = blockIdx.y;
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BT) tw[i] = twg[i];
  const int l = tid % LPB, t = tid / LPB;
  const bool act = l < li.len;
  constexpr int64_t npos = (int64_t)N * NH;
  const int64_t pos = li.pos0 + l;
  cplx<T>* sp = a.spec + cd.off_spec + (int64_t)rhs * 3 * cd.nspec;
  // input: the residual's spectrum when the plan keeps one (out of place), else spec itself
  const cplx<T>* spi = (a.rspec != nullptr ? a.rspec : a.spec) + cd.off_spec + (int64_t)rhs * 3 * cd.nspec;
  // the three components' columns are loaded before the first exchange (all loads in flight together)
  cplx<T> v[3][N1];
  if (act && t < N2) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) v[c][j1] = spi[(int64_t)c * cd.nspec + (int64_t)(t + N2 * j1) * npos + pos];
  }
  __syncthreads();
  cplx<T> Xc[3][N2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // forward step 1: length N1 over x = t + N2 j1, twiddled into the exchange buffer
    if (act && t < N2) {
      Dft<T, N1, -1>::run(v[c]);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) xch[(k1 * (N2 + 1) + t) * LPB + l] = cmulw(v[c][k1], tw[t * k1], false);
    }
    __syncthreads();
    // forward step 2: length N2, outputs kx = t + N1 k2
    if (act && t < N1) {
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) Xc[c][j2] = xch[(t * (N2 + 1) + j2) * LPB + l];
      Dft<T, N2, -1>::run(Xc[c]);
    }
    __syncthreads();
  }
  if (act && t < N1) {
    // z = Kinv r at kx = t + N1 k2 (symbol component-planar: [6][nx][ny][nzh], divided by N^3)
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int64_t o = (int64_t)(t + N1 * k2) * npos + pos;
      constexpr int64_t ss = (int64_t)N * npos;
      const T k00 = symp[o], k11 = symp[ss + o], k22 = symp[2 * ss + o];
      const T k12 = symp[3 * ss + o], k02 = symp[4 * ss + o], k01 = symp[5 * ss + o];
      const cplx<T> r0 = Xc[0][k2], r1 = Xc[1][k2], r2 = Xc[2][k2];
      Xc[0][k2] = {k00 * r0.re + k01 * r1.re + k02 * r2.re, k00 * r0.im + k01 * r1.im + k02 * r2.im};
      Xc[1][k2] = {k01 * r0.re + k11 * r1.re + k12 * r2.re, k01 * r0.im + k11 * r1.im + k12 * r2.im};
      Xc[2][k2] = {k02 * r0.re + k12 * r1.re + k22 * r2.re, k02 * r0.im + k12 * r1.im + k22 * r2.im};
      if (DOTB) {
        const double d = (double)r0.re * (double)Xc[0][k2].re + (double)r0.im * (double)Xc[0][k2].im +
                         (double)r1.re * (double)Xc[1][k2].re + (double)r1.im * (double)Xc[1][k2].im +
                         (double)r2.re * (double)Xc[2][k2].re + (double)r2.im * (double)Xc[2][k2].im;
        const int kz = (int)(pos % NH);
        dacc += (kz == 0 || 2 * kz == N) ? d : 2.0 * d;
      }
    }
  }
  if (DOTB) {
    const double tsum = block_sum(dacc, red);
    if (tid == 0) a.part_bdot[(int64_t)rhs * a.nlines + a.line_base + blockIdx.x] = tsum;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // inverse step 1: in-register length N2 at kx = t + N1 k2; twiddle e^{+2 pi i t k / N}
    if (act && t < N1) {
      Dft<T, N2, 1>::run(Xc[c]);
#pragma unroll
      for (int k1 = 0; k1 < N2; ++k1) xch[(k1 * (N1 + 1) + t) * LPB + l] = cmulw(Xc[c][k1], tw[t * k1], true);
    }
    __syncthreads();
    // inverse step 2: thread t < N2 gathers over t' < N1, outputs x = t + N2 k
    if (act && t < N2) {
      cplx<T> w[N1];
#pragma unroll
      for (int j = 0; j < N1; ++j) w[j] = xch[(t * (N1 + 1) + j) * LPB + l];
      Dft<T, N1, 1>::run(w);
#pragma unroll
      for (int k = 0; k < N1; ++k) sp[(int64_t)c * cd.nspec + (int64_t)(t + N2 * k) * npos + pos] = w[k];
    }
    if (c < 2) __syncthreads();
  }
}

// Launchers ------------------------------------------------------------------------------------
template <typename T, int N>
size_t fast_lds_plane() {
  using C = FastCfg<T, N>;
  return ((size_t)N + (size_t)N * C::NH + (C::XIN ? 0 : (size_t)C::LPR * C::XS) + N) * sizeof(cplx<T>);
}
template <typename T, int N>
size_t fast_lds_line(int lpb) {
  using C = FastCfg<T, N>;
  return ((size_t)N + (size_t)C::XS * (lpb > 0 ? lpb : C::LPB)) * sizeof(cplx<T>);
}

// pass B positions per block other than the default that the FP32 passes offer (half and double of it)
template <int N>
constexpr bool pass_b_alt_ok(int lpb) {
  return lpb == FastCfg<float, N>::LPB / 2 ||
         (lpb == 2 * FastCfg<float, N>::LPB && FastCfg<float, N>::L * lpb <= 1024);
}

// update: 0 none, 2 r -= alpha Ap and x += alpha p in pass A; 1 (r only) and 4 (spectral residual) on the
// staged pass only;
// dot: 0 none, 1 r . z in pass C, 2 r . z in pass B (spectral), 3 as 2 and pass C forms the new search
// direction, 4 as 3 with the paired iterate step of odd iterations (staged pass only, see prec_c_staged)
template <typename T, int N>
static hipError_t launch_fast_n(const PrecArgs<T>& a, int nplanes, int nlines, const cplx<T>* tw, const T* symp,
                                int update, int dot, int vec, int lpb, int ppa, hipStream_t st) {
  using C = FastCfg<T, N>;
  if (lpb == C::LPB) lpb = 0;
  if (lpb != 0 && (sizeof(T) != 4 || !pass_b_alt_ok<N>(lpb))) return hipErrorInvalidValue;
  const size_t lp = fast_lds_plane<T, N>(), ll = fast_lds_line<T, N>(lpb);
  dim3 gp(nplanes, a.m, 1), gl(nlines, a.m, 1);
  auto big = [](const void* f, size_t b) {
    if (b > 65536) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
  };
#define NLFH_A(KERN, U)                                                                 \
  big((const void*)KERN<T, N, U>, lp);                                                \
  hipLaunchKernelGGL((KERN<T, N, U>), gp, dim3(C::BT), lp, st, a, tw);
#define NLFH_B1(D, X)                                                                   \
  big((const void*)prec_b_fast<T, N, D, X>, ll);                                      \
  hipLaunchKernelGGL((prec_b_fast<T, N, D, X>), gl, dim3(pass_b_threads<T, N, X>()), ll, st, a, tw, symp);
#define NLFH_B(D)                                                                       \
  if constexpr (sizeof(T) == 4) {                                                     \
    if (lpb == C::LPB / 2) { NLFH_B1(D, C::LPB / 2) }                                  \
    else if (lpb == 2 * C::LPB) { NLFH_B1(D, (C::L * 2 * C::LPB <= 1024 ? 2 * C::LPB : C::LPB)) }  \
    else { NLFH_B1(D, 0) }                                                           \
  } else {                                                                            \
    NLFH_B1(D, 0)                                                                     \
  }
  if constexpr (StageCfg<T, N>::fits) {
    if (vec) {
      using SC = StageCfg<T, N>;
      const size_t ls = SC::lds_bytes;
      const dim3 bs(SC::BT);
#define NLFH_AS(U)                                                                      \
  big((const void*)prec_a_staged<T, N, U>, ls);                                       \
  hipLaunchKernelGGL((prec_a_staged<T, N, U>), gp, bs, ls, st, a, tw);
#define NLFH_CS(D, DIRV)                                                                \
  big((const void*)prec_c_staged<T, N, D, DIRV>, ls);                                 \
  hipLaunchKernelGGL((prec_c_staged<T, N, D, DIRV>), gp, bs, ls, st, a, tw);
      bool two = false;
      if constexpr (sizeof(T) == 4 && !SC::IP) {
        if (update == 4 && ppa == 2) {
          two = true;
          big((const void*)prec_a_staged<T, N, 4, 2>, ls);
          hipLaunchKernelGGL((prec_a_staged<T, N, 4, 2>), dim3((nplanes + 1) / 2, a.m, 1), bs, ls, st, a, tw);
        }
      }
      if (two) {}
      else if (update == 4) { NLFH_AS(4) }
      else if (update == 1) { NLFH_AS(1) }
      else if (update) { NLFH_AS(2) }
      else { NLFH_AS(0) }
      if (dot >= 2) { NLFH_B(true) }
      else { NLFH_B(false) }
      if (dot == 1) { NLFH_CS(true, 0) }
      else if (dot == 3) { NLFH_CS(false, 1) }
      else if (dot == 4) { NLFH_CS(false, 2) }
      else { NLFH_CS(false, 0) }
#undef NLFH_AS
#undef NLFH_CS
      return hipGetLastError();
    }
  }
  // r-only and spectral updates, the direction pass and a kept residual spectrum: staged passes only
  // (and alpha formed from the matvec's partials: a.part_mv with an update)
  if (update == 1 || update == 4 || dot >= 3 || a.rspec != nullptr || (update && a.part_mv != nullptr))
    return hipErrorInvalidValue;
  if (update) { NLFH_A(prec_a_fast, true) }
  else { NLFH_A(prec_a_fast, false) }
  if (dot == 2) { NLFH_B(true) }
  else { NLFH_B(false) }
#undef NLFH_A
#undef NLFH_B
#undef NLFH_B1
  if (dot == 1) {
    big((const void*)prec_c_fast<T, N, true>, lp);
    hipLaunchKernelGGL((prec_c_fast<T, N, true>), gp, dim3(C::BT), lp, st, a, tw);
  } else {
    big((const void*)prec_c_fast<T, N, false>, lp);
    hipLaunchKernelGGL((prec_c_fast<T, N, false>), gp, dim3(C::BT), lp, st, a, tw);
  }
  return hipGetLastError();
}

bool fast_fft_size(int n) { return n == 16 || n == 32 || n == 48 || n == 64 || n == 128; }

bool fast_staged(int n, int elem_bytes) {
  switch (n) {
    case 16: return elem_bytes == 8 ? StageCfg<double, 16>::fits : StageCfg<float, 16>::fits;
    case 32: return elem_bytes == 8 ? StageCfg<double, 32>::fits : StageCfg<float, 32>::fits;
    case 48: return elem_bytes == 8 ? StageCfg<double, 48>::fits : StageCfg<float, 48>::fits;
    case 64: return elem_bytes == 8 ? StageCfg<double, 64>::fits : StageCfg<float, 64>::fits;
    case 128: return elem_bytes == 8 ? StageCfg<double, 128>::fits : StageCfg<float, 128>::fits;
    default: return false;
  }
}

int fast_line_len(int n, int elem_bytes) {
  switch (n) {
    case 16: return elem_bytes == 8 ? FastCfg<double, 16>::LPB : FastCfg<float, 16>::LPB;
    case 32: return elem_bytes == 8 ? FastCfg<double, 32>::LPB : FastCfg<float, 32>::LPB;
    case 48: return elem_bytes == 8 ? FastCfg<double, 48>::LPB : FastCfg<float, 48>::LPB;
    case 128: return elem_bytes == 8 ? FastCfg<double, 128>::LPB : FastCfg<float, 128>::LPB;
    default: return elem_bytes == 8 ? FastCfg<double, 64>::LPB : FastCfg<float, 64>::LPB;
  }
}

bool fast_two_planes_ok(int n) {
  switch (n) {
    case 16: return StageCfg<float, 16>::fits && !StageCfg<float, 16>::IP;
    case 32: return StageCfg<float, 32>::fits && !StageCfg<float, 32>::IP;
    case 48: return StageCfg<float, 48>::fits && !StageCfg<float, 48>::IP;
    case 64: return StageCfg<float, 64>::fits && !StageCfg<float, 64>::IP;
    case 128: return StageCfg<float, 128>::fits && !StageCfg<float, 128>::IP;
    default: return false;
  }
}

bool fast_pass_b_ok(int n, int lpb) {
  if (lpb == 0) return true;
  switch (n) {
    case 16: return lpb == FastCfg<float, 16>::LPB || pass_b_alt_ok<16>(lpb);
    case 32: return lpb == FastCfg<float, 32>::LPB || pass_b_alt_ok<32>(lpb);
    case 48: return lpb == FastCfg<float, 48>::LPB || pass_b_alt_ok<48>(lpb);
    case 64: return lpb == FastCfg<float, 64>::LPB || pass_b_alt_ok<64>(lpb);
    case 128: return lpb == FastCfg<float, 128>::LPB || pass_b_alt_ok<128>(lpb);
    default: return false;
  }
}

template <typename T>
hipError_t launch_prec_fast(int n, const PrecArgs<T>& a, int nplanes, int nlines, const cplx<T>* tw, const T* symp,
                            int update, int dot, int vec, int lpb, int ppa, hipStream_t st) {
  if (ppa != 1 && ppa != 2) return hipErrorInvalidValue;
  switch (n) {
    case 16: return launch_fast_n<T, 16>(a, nplanes, nlines, tw, symp, update, dot, vec, lpb, ppa, st);
    case 32: return launch_fast_n<T, 32>(a, nplanes, nlines, tw, symp, update, dot, vec, lpb, ppa, st);
    case 48: return launch_fast_n<T, 48>(a, nplanes, nlines, tw, symp, update, dot, vec, lpb, ppa, st);
    case 64: return launch_fast_n<T, 64>(a, nplanes, nlines, tw, symp, update, dot, vec, lpb, ppa, st);
    case 128: return launch_fast_n<T, 128>(a, nplanes, nlines, tw, symp, update, dot, vec, lpb, ppa, st);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_prec_fast<double>(int, const PrecArgs<double>&, int, int, const cplx<double>*,
                                             const double*, int, int, int, int, int, hipStream_t);
template hipError_t launch_prec_fast<float>(int, const PrecArgs<float>&, int, int, const cplx<float>*, const float*,
                                            int, int, int, int, int, hipStream_t);

}  // namespace nlfh
