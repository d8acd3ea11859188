// Matrix-free Q1 elasticity operator K(s) u (+ the macroscopic-strain load), gfx950.
//
// Restates nlfiber/linear.py:143-149 (apply_K = scatter(s * (gather(u) @ Ke)), load =
// scatter(s * E.G)) without the 24x24 element matrix. With unnormalised Walsh sums
// h = W u_e (W = Wz (x) Wy (x) Wx, two-point sum/difference per axis) the unit-cube Q1 stiffness
// with 2x2x2 Gauss points is Ke = W^T L W with a sparse L: the constant-strain block D/16,
// the three bilinear modes coupled in pairs/triples with weights (lambda, mu)/48 and the trilinear
// mode (lambda + 4 mu)/144 (derived from the quadrature-orthogonality of the Walsh modes; checked
// against the dense Ke to round-off in the tests). The butterflies are shared by neighbouring
// elements, so one element costs about 140 VALU operations per right-hand side instead of 1152.
//
// Tiling: a block owns a y-tile of TY node rows and all nz of one cell, and marches over an
// x-chunk of XC planes (one extra plane of warm-up). A row of nz nodes is LZ lanes x ZPT
// consecutive z values in a slot of LZS lanes (a power of two <= 64, so rows never straddle a
// wavefront); the x
// butterflies live in registers across the march, z neighbours come from the adjacent lane
// (__shfl), y neighbours through LDS (three small buffers, two barriers per plane). The tile
// loads TY + 2 rows and computes TY + 1 element rows (one halo row each side). When nz / ZPT is
// not a power of two the row is padded to the next power of two lanes (idle lanes), e.g. 48
// lanes in a slot of 64.
#include <type_traits>

#include "nlfh_kernels.h"

namespace nlfh {


template <typename T>
__device__ __forceinline__ void modal_op(const T (&h)[3][8], T (&g)[3][8], const T l1, const T m1,
                                         const T* e4) {
  T exx = h[0][1], eyy = h[1][2], ezz = h[2][4];
  T gyz = h[1][4] + h[2][2], gxz = h[0][4] + h[2][1], gxy = h[0][2] + h[1][1];
  if (e4) {
    exx += e4[0]; eyy += e4[1]; ezz += e4[2];
    gyz += e4[3]; gxz += e4[4]; gxy += e4[5];
  }
  const T tr = exx + eyy + ezz;
  const T ltr = l1 * tr;
  const T m2x = m1 + m1;
  g[0][1] = ltr + m2x * exx;
  g[1][2] = ltr + m2x * eyy;
  g[2][4] = ltr + m2x * ezz;
  const T tyz = m1 * gyz, txz = m1 * gxz, txy = m1 * gxy;
  g[1][4] = tyz; g[2][2] = tyz;
  g[0][4] = txz; g[2][1] = txz;
  g[0][2] = txy; g[1][1] = txy;
  const T third = T(1) / T(3);
  const T l2 = l1 * third, m2 = m1 * third;
  const T a2 = l2 + T(3) * m2;
  g[0][3] = a2 * h[0][3] + l2 * h[2][6];
  g[2][6] = a2 * h[2][6] + l2 * h[0][3];
  g[1][3] = a2 * h[1][3] + l2 * h[2][5];
  g[2][5] = a2 * h[2][5] + l2 * h[1][3];
  g[0][5] = a2 * h[0][5] + l2 * h[1][6];
  g[1][6] = a2 * h[1][6] + l2 * h[0][5];
  const T t = h[2][3] + h[1][5] + h[0][6];
  g[2][3] = m2 * (t + h[2][3]);
  g[1][5] = m2 * (t + h[1][5]);
  g[0][6] = m2 * (t + h[0][6]);
  const T kap = (l1 + T(4) * m1) * (T(1) / T(9));
  g[0][7] = kap * h[0][7];
  g[1][7] = kap * h[1][7];
  g[2][7] = kap * h[2][7];
  g[0][0] = T(0); g[1][0] = T(0); g[2][0] = T(0);
}

// z neighbours by the kernel's DPP mode: 1, a row is a whole wavefront: a wave rotation (lane i <- i+1 with
// wave_rol:1, i <- i-1 with wave_ror:1, checked on gfx950); 2, a row is one 16-lane DPP row (nz / zpt = 16): a row
// rotation (lane i <- i+1 of its row with row_ror:15, i <- i-1 with row_ror:1); 0, ds_bpermute.
// A wave rotation reads a valid source lane for every lane, so the instruction's `old` operand is
// never used: mov_dpp leaves it undefined, which spares a register copy per rotation and lets the
// compiler fold the rotation into the adds and subtracts that consume it.
template <int CTRL>
__device__ __forceinline__ double dpp_rot(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_rot(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
// Marks the registers of a prefetch as read here, on every path, with no instruction: the wait for the
// load lands at this point (one step after the load was issued) and the registers are settled
// afterwards, so the compiler need not wait early on the paths where the consumer is skipped and the
// register is reused.
template <typename T, int K>
__device__ __forceinline__ void settle(const T (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) asm volatile("" ::"v"(v[k]));
}

// s_waitcnt immediate (gfx9 encoding): vmcnt = 0, expcnt and lgkmcnt at their maxima (no wait)
constexpr int kWaitVmcnt0 = 0x0F70;
constexpr int kDppWaveRol1 = 0x134;  // lane i receives lane i + 1 (mod 64)
constexpr int kDppWaveRor1 = 0x13C;  // lane i receives lane i - 1 (mod 64)
constexpr int kDppRowRor15 = 0x12F;  // lane i receives lane i + 1 of its 16-lane row (mod 16)
constexpr int kDppRowRor1 = 0x121;   // lane i receives lane i - 1 of its 16-lane row (mod 16)

// z butterflies of the last z value of a lane with the first of the next lane, one component, FP32, full
// wavefront rows: h[mm] = next(a[mm]) + b[mm] (mm = 1..3; the sum of mode 0 is never used by the element
// operator) and h[mm + 4] = next(a[mm]) - b[mm] (mm = 0..3), with the rotation (lane i <- lane i + 1)
// applied to the first operand of each add/sub instead of a separate move. The leading s_nop covers the
// two wait states that a DPP read of a register written by the previous VALU instruction needs.
__device__ __forceinline__ void z_next_butterflies(const float (&a)[4], const float (&b)[4], float (&h)[8]) {
  asm("s_nop 1\n\t"
      "v_sub_f32_dpp %3, %7, %11 wave_rol:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_add_f32_dpp %0, %8, %12 wave_rol:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_sub_f32_dpp %4, %8, %12 wave_rol:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_add_f32_dpp %1, %9, %13 wave_rol:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_sub_f32_dpp %5, %9, %13 wave_rol:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_add_f32_dpp %2, %10, %14 wave_rol:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_sub_f32_dpp %6, %10, %14 wave_rol:1 row_mask:0xf bank_mask:0xf"
      : "=&v"(h[1]), "=&v"(h[2]), "=&v"(h[3]), "=&v"(h[4]), "=&v"(h[5]), "=&v"(h[6]), "=&v"(h[7])
      : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(b[0]), "v"(b[1]), "v"(b[2]), "v"(b[3]));
  h[0] = 0.0f;  // unused
}

template <typename T, int DPP>
__device__ __forceinline__ T from_next(T v, int src) {
  if constexpr (DPP == 1) return dpp_rot<kDppWaveRol1>(v);
  else if constexpr (DPP == 2) return dpp_rot<kDppRowRor15>(v);
  else return __shfl(v, src, 64);
}
template <typename T, int DPP>
__device__ __forceinline__ T from_prev(T v, int src) {
  if constexpr (DPP == 1) return dpp_rot<kDppWaveRor1>(v);
  else if constexpr (DPP == 2) return dpp_rot<kDppRowRor1>(v);
  else return __shfl(v, src, 64);
}

// Element i of one component field of one cell, addressed as a 32-bit byte offset from the field's
// wave-uniform base, so that loads and stores use the scalar-base + 32-bit vector-offset form with no
// 64-bit address arithmetic per access. The largest accepted cell (512 x 512 x 256 voxels) has
// 512 MiB per FP64 component field, so the byte offset fits in 32 bits.
template <typename T>
__device__ __forceinline__ const T& at(const T* base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T& at(T* base, unsigned i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T));
}

// One block: y-tile rows [y0, y0 + ty) x all z, x-planes [x0, x0 + xc), one right-hand side.
// Element layer L_s = x0 - 1 + s (s = 0..xc) lies between node planes L_s and L_s + 1. Step s
// (one barrier) does: (a) finish layer L_{s-1}: add the row-above half of its y-reverse from LDS,
// reverse x and store node plane L_{s-1}; (b) layer L_s: y, z butterflies, modal operator, z and y
// reverse, publish the row-above half in LDS; (c) x butterflies of layer L_{s+1} from the prefetched
// plane into the other LDS buffer. LDS: two S/D buffers, selected by the parity of s, which is a
// compile-time constant (the step loop is unrolled by two), and one y-reverse buffer, which part (b)
// rewrites after a barrier that follows part (a)'s reads: three buffers instead of four let four blocks of
// a 6-row tile share a CU (8 wavefronts per SIMD instead of 6; the standalone FP32 operator on 8 x 64^3
// cells 106 -> 95 us).
// Addressing: every global access is a wave-uniform base (cell, right-hand side, component, x-plane)
// plus a 32-bit per-lane offset inside the plane; with a row per wavefront (DPP mode 1) the row index is
// wave-uniform too. In DPP modes 1 and 2 a row is a whole wavefront or a whole 16-lane DPP row, and the conditions
// that mask the element work (active, erow) depend on the row alone, so a rotation only reads lanes of its own
// active row.
// S64: the stiffness is read in FP64 (a.s64) and rounded to T as it is loaded (the standalone FP32
// operator, whose caller holds s in FP64; the solver converts s once per solve and reads a.s).
// CHEB (STRAIN only): the owned node values become one step of the Chebyshev-Jacobi iteration for K u = -f,
// u + c1 (u - x_prev) - c2 dinv (K u + f); x_prev and dinv of the plane stored at step s are loaded at step s - 1.
template <typename T, int ZPT, bool STRAIN, bool PUPDATE, bool DOT, int DPP, bool S64 = false, bool CHEB = false>
__global__ void __launch_bounds__(512) matvec_kernel(const MatvecArgs<T> a) {
  static_assert(!CHEB || (STRAIN && !PUPDATE && !DOT && !S64), "the Chebyshev step is a strain-term launch");
  extern __shared__ double smem_d[];
  T* const lds = reinterpret_cast<T*>(smem_d);
  __shared__ double red[8];

  const MvItem it = a.items[blockIdx.x];
  if (a.done != nullptr && a.done[it.cell]) return;
  const CellDesc cd = a.cells[it.cell];
  const int rhs = blockIdx.y;
  const int nx = cd.nx, ny = cd.ny, nz = cd.nz;
  const int LZ = DPP == 1 ? 64 : (DPP == 2 ? 16 : a.LZ);
  const int LZS = DPP == 1 ? 64 : (DPP == 2 ? 16 : a.LZS);
  const int ty = it.ty;
  const int RB = ty + 2;
  const int tid = threadIdx.x;
  const int lz = DPP == 1 ? (tid & 63) : (DPP == 2 ? (tid & 15) : tid % LZS);
  const int r = DPP == 1 ? __builtin_amdgcn_readfirstlane(tid >> 6) : (DPP == 2 ? (tid >> 4) : tid / LZS);
  const bool active = r < RB && lz < LZ;
  const bool erow = r <= ty;  // computes element row yg
  const bool owns = active && r >= 1 && r <= ty;
  int yg = it.y0 - 1 + r;
  yg = yg < 0 ? yg + ny : yg;
  yg = yg >= ny ? yg - ny : yg;
  const int z0 = lz * ZPT;
  const int lane = tid & 63;
  const int src_next = (lane - lz) + (lz + 1 == LZ ? 0 : lz + 1);
  const int src_prev = (lane - lz) + (lz == 0 ? LZ - 1 : lz - 1);

  const int64_t nvox = cd.nvox;
  const int plane = ny * nz;
  const unsigned roff = (unsigned)(yg * nz + z0);  // offset of this lane's nodes inside a plane
  const int64_t fbase = cd.off_field + (int64_t)rhs * 3 * nvox;
  const T* Uc[3];
  const T* Pc[3];
  T* PNc[3];
  T* Oc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Uc[c] = a.u ? a.u + fbase + c * nvox : nullptr;
    Pc[c] = PUPDATE ? a.p_old + fbase + c * nvox : nullptr;
    PNc[c] = PUPDATE ? a.p_new + fbase + c * nvox : nullptr;
    Oc[c] = a.out + fbase + c * nvox;
  }
  const T* Xp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) Xp[c] = CHEB ? a.x_prev + fbase + c * nvox : nullptr;
  const T* __restrict__ DI = CHEB ? a.dinv + cd.off_vox : nullptr;
  const T c1 = CHEB ? a.c1 : T(0), c2 = CHEB ? a.c2 : T(0);
  const T* __restrict__ S = S64 ? nullptr : a.s + cd.off_vox;
  const double* __restrict__ S8 = S64 ? a.s64 + cd.off_vox : nullptr;
  auto load_s = [&](unsigned i) -> T {
    if constexpr (S64) return (T)at(S8, i);
    else return at(S, i);
  };
  T beta = T(0);
  if (PUPDATE) beta = (T)a.beta[it.cell * a.m + rhs];
  T e4[6];
  if (STRAIN) {
#pragma unroll
    for (int j = 0; j < 6; ++j) e4[j] = (T)(4.0 * a.strain[rhs * 6 + j]);
  }
  const T lam16 = a.lam16, mu16 = a.mu16, scale_out = a.scale_out;

  const int buf_sz = RB * 6 * ZPT * LZ;  // one LDS buffer
  const int lbase = r * 6 * ZPT * LZ + lz;
  const int rowstep = 6 * ZPT * LZ;
#define LOFF(f, q) (((f) * ZPT + (q)) * LZ)

  // x-plane index -> offset of the plane (x in [x0 - 2, x0 + xc + 1], wrapped periodically)
  auto plane_off = [&](int x) -> unsigned {
    x = x < 0 ? x + nx : x;
    x = x < 0 ? x + nx : x;
    x = x >= nx ? x - nx : x;
    x = x >= nx ? x - nx : x;
    return (unsigned)(x * plane);
  };
  // Inactive lanes (idle lanes of a padded row, and rows r >= RB of a block that has more threads than
  // the tile's rows: its thread count is rounded up to whole wavefronts and sized for the largest tile)
  // take part in the step prefetch, which is issued on every path: they read their plane's first row.
  // Their own row index is not a valid row (y0 - 1 + r may exceed 2 ny, beyond the single wrap above).
  const unsigned roff_ld = active ? roff : 0u;
  const bool has_u = !STRAIN || a.u != nullptr;  // u may be null only in load launches (host-checked)
  // loads only: the values are combined (u + beta p) where they are consumed, so that a plane
  // prefetched one step ahead stays in flight across the step's barrier
  auto load_row = [&](unsigned po, T (&vu)[3][ZPT], T (&vp)[3][ZPT]) {
    const unsigned o = po + roff_ld;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int q = 0; q < ZPT; ++q) {
        vu[c][q] = has_u ? at(Uc[c], o + q) : T(0);
        if (PUPDATE) vp[c][q] = at(Pc[c], o + q);
      }
    }
  };
  // p_new of an owned node plane
  auto store_dir = [&](unsigned o, const T (&un)[3][ZPT]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < ZPT; ++q) at(PNc[c], o + q) = un[c][q];
  };
  auto combine = [&](const T (&vu)[3][ZPT], const T (&vp)[3][ZPT], T (&v)[3][ZPT]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < ZPT; ++q) v[c][q] = PUPDATE ? vu[c][q] + beta * vp[c][q] : vu[c][q];
  };
  const int x0 = it.x0, xc = it.xc;

  T up[3][ZPT], carry[3][ZPT];
  // one step ahead, by the parity of the step that loads them (fixed registers, no copies of values
  // in flight): the next x-plane (u and p) and the stiffness of the next element layer
  T nxt_u[2][3][ZPT], nxt_p[2][3][ZPT], sv[2][ZPT];
  T pc[2][3][ZPT];  // node values of the planes awaiting output (by layer parity)
  T xq[2][3][ZPT], dq[2][ZPT];  // CHEB: previous iterate and 1 / diagonal of the next plane to store
  T Flo[3][2][ZPT];  // own half of the y-reverse of the last element layer
  double dot = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < ZPT; ++q) {
      carry[c][q] = T(0);
      Flo[c][0][q] = T(0);
      Flo[c][1][q] = T(0);
      pc[0][c][q] = T(0);
      pc[1][c][q] = T(0);
      xq[0][c][q] = xq[1][c][q] = T(0);
      if (c == 0) dq[0][q] = dq[1][q] = T(0);
    }

  // x stage of layer L_t (parity PT = t & 1) from planes `up` (L_t) and `un` (L_t + 1)
  auto x_stage = [&](auto ptag, T (&un)[3][ZPT]) {
    constexpr int PT = decltype(ptag)::value;
    T* SD = lds + PT * buf_sz + lbase;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < ZPT; ++q) {
        SD[LOFF(c, q)] = un[c][q] + up[c][q];
        SD[LOFF(3 + c, q)] = un[c][q] - up[c][q];
        pc[PT][c][q] = up[c][q];
        up[c][q] = un[c][q];
      }
  };

  // prologue: layer L_0 from planes x0-1, x0; prefetch x0+1
  if (active) {
    T un[3][ZPT], lu[3][ZPT], lp[3][ZPT];
    load_row(plane_off(x0 - 1), lu, lp);
    combine(lu, lp, up);
    const unsigned p0 = plane_off(x0);
    load_row(p0, lu, lp);
    combine(lu, lp, un);
    if (xc >= 1) load_row(plane_off(x0 + 1), nxt_u[1], nxt_p[1]);
    {
      const unsigned so = plane_off(x0 - 1) + roff_ld;
#pragma unroll
      for (int q = 0; q < ZPT; ++q) sv[1][q] = load_s(so + q);
    }
    if (PUPDATE && owns && xc >= 1) store_dir(p0 + roff, un);
    x_stage(std::integral_constant<int, 0>{}, un);
  }
  // Wave-uniform offsets of the planes the step loop touches, advanced by one plane per step with a single
  // wrap instead of recomputed from the plane index (two wraps each way) three times per step: the x-plane
  // prefetched for the next step, x0 + min(s + 2, xc); the stiffness layer, x0 + min(s, xc - 1); the node plane
  // stored, x0 - 2 + s; CHEB: the plane of x_prev and 1 / diag, x0 - 1 + min(s, xc).
  const unsigned cell_span = (unsigned)nx * (unsigned)plane;
  unsigned off_ld = plane_off(x0 + (2 < xc ? 2 : xc));
  unsigned off_sv = plane_off(x0);
  unsigned off_st = plane_off(x0 - 2);
  unsigned off_xq = plane_off(x0 - 1);
  auto advance = [&](unsigned& o) {
    o += (unsigned)plane;
    o = o >= cell_span ? o - cell_span : o;
  };
  // Drain the prologue's loads here (they were issued just after the ones x_stage waited for): the
  // step loop's entry then carries no outstanding load, and its steady-state waits cover only the
  // plane prefetched one step earlier instead of the whole queue at the top of every step.
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();

  // one step; returns true when the march is over (s > xc after part (a))
  auto step = [&](int s, auto ptag) -> bool {
    constexpr int PS = decltype(ptag)::value;  // s & 1
    constexpr int PQ = 1 - PS;                 // (s - 1) & 1 == (s + 1) & 1
    // prefetch for the next step, issued on every path (the last steps repeat the chunk's last plane,
    // a cache hit): x-plane x0 + s + 2 (consumed by part (c) of step s + 1) and the stiffness of
    // element layer L_{s+1} (part (b) of step s + 1)
    {
      load_row(off_ld, nxt_u[PS], nxt_p[PS]);
      const unsigned so = off_sv + roff_ld;
#pragma unroll
      for (int q = 0; q < ZPT; ++q) sv[PS][q] = load_s(so + q);
      if constexpr (CHEB) {
        // node plane x0 - 1 + s, stored by part (a) of step s + 1 (the last steps repeat plane x0 + xc - 1)
        const unsigned xo = off_xq + roff_ld;
#pragma unroll
        for (int q = 0; q < ZPT; ++q) {
#pragma unroll
          for (int c = 0; c < 3; ++c) xq[PS][c][q] = at(Xp[c], xo + q);
          dq[PS][q] = at(DI, xo + q);
        }
      }
    }
    // (a) finish layer L_{s-1}: node plane L_{s-1} = x0 - 2 + s is stored for s >= 2
    if (s >= 1 && owns) {
      const T* UYp = lds + 2 * buf_sz + lbase - rowstep;
      const unsigned oo = off_st + roff;
      T dsum = T(0);  // float32: the step's p . Ap terms summed in float, added to the double total once
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < ZPT; ++q) {
          const T A = Flo[c][0][q] + UYp[LOFF(2 * c, q)];
          const T B = Flo[c][1][q] + UYp[LOFF(2 * c + 1, q)];
          const T lo = A - B, hi = A + B;
          if (s >= 2) {
            T v = carry[c][q] + lo;
            if (STRAIN) v *= scale_out;  // only the load launches scale (by -1); the others pass 1
            if constexpr (CHEB) {
              const T xk = pc[PQ][c][q];
              v = xk + c1 * (xk - xq[PQ][c][q]) - c2 * dq[PQ][q] * v;
            }
            at(Oc[c], oo + q) = v;
            if (DOT) {
              if constexpr (sizeof(T) == 8) dot += pc[PQ][c][q] * v;
              else dsum += pc[PQ][c][q] * v;
            }
          }
          carry[c][q] = hi;
        }
      if (DOT && sizeof(T) == 4 && s >= 2) dot += (double)dsum;
    }
    // the y-reverse exchange is a single buffer: part (b) below overwrites what part (a) read
    __syncthreads();
    if (s > xc) return true;
    // (b) element layer L_s
    settle(sv[PQ]);
    if constexpr (CHEB) {  // consumed by part (a) above
#pragma unroll
      for (int c = 0; c < 3; ++c) settle(xq[PQ][c]);
      settle(dq[PQ]);
    }
    if (active && erow) {
      const T* SD = lds + PS * buf_sz + lbase;
      T* UY = lds + 2 * buf_sz + lbase;
      T Y[3][4][ZPT];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < ZPT; ++q) {
          const T s0 = SD[LOFF(c, q)], d0 = SD[LOFF(3 + c, q)];
          const T s1 = SD[rowstep + LOFF(c, q)], d1 = SD[rowstep + LOFF(3 + c, q)];
          Y[c][0][q] = s1 + s0;
          Y[c][2][q] = s1 - s0;
          Y[c][1][q] = d1 + d0;
          Y[c][3][q] = d1 - d0;
        }
      // FP32 full-wavefront rows fold the rotation into the butterflies (z_next_butterflies)
      constexpr bool kFoldNext = DPP == 1 && std::is_same<T, float>::value;
      T Yn[3][4];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) Yn[c][mm] = kFoldNext ? T(0) : from_next<T, DPP>(Y[c][mm][0], src_next);
      T sv_cur[ZPT];
#pragma unroll
      for (int q = 0; q < ZPT; ++q) sv_cur[q] = sv[PQ][q];
      T Fz[3][4][ZPT];
      T hi_last[3][4];
#pragma unroll
      for (int q = 0; q < ZPT; ++q) {
        T h[3][8], g[3][8];
        if constexpr (kFoldNext) {
          if (q + 1 == ZPT) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float a[4] = {Y[c][0][0], Y[c][1][0], Y[c][2][0], Y[c][3][0]};
              const float b[4] = {Y[c][0][q], Y[c][1][q], Y[c][2][q], Y[c][3][q]};
              z_next_butterflies(a, b, h[c]);
            }
          }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) {
            if (kFoldNext && q + 1 == ZPT) continue;
            const T yn = (q + 1 < ZPT) ? Y[c][mm][(q + 1) % ZPT] : Yn[c][mm];
            h[c][mm] = yn + Y[c][mm][q];
            h[c][mm + 4] = yn - Y[c][mm][q];
          }
        const T sv = sv_cur[q];
        modal_op<T>(h, g, sv * lam16, sv * mu16, STRAIN ? e4 : nullptr);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) {
            // mode 0 of the element forces is zero (a constant field has no energy)
            const T lo = mm == 0 ? -g[c][4] : g[c][mm] - g[c][mm + 4];
            const T hi = mm == 0 ? g[c][4] : g[c][mm] + g[c][mm + 4];
            if (q == 0) Fz[c][mm][0] = lo;
            else Fz[c][mm][q] += lo;
            if (q + 1 < ZPT) Fz[c][mm][(q + 1) % ZPT] = hi;
            else hi_last[c][mm] = hi;
          }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) Fz[c][mm][0] += from_prev<T, DPP>(hi_last[c][mm], src_prev);
      // y reverse: the row-above half goes to row r + 1 through LDS, the own half waits in Flo
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < ZPT; ++q) {
          const T A0 = Fz[c][0][q], B0 = Fz[c][2][q];
          const T A1 = Fz[c][1][q], B1 = Fz[c][3][q];
          UY[LOFF(2 * c, q)] = A0 + B0;
          UY[LOFF(2 * c + 1, q)] = A1 + B1;
          Flo[c][0][q] = A0 - B0;
          Flo[c][1][q] = A1 - B1;
        }
    }
    // (c) x stage of layer L_{s+1} from planes x0 + s and x0 + s + 1 (prefetched)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      settle(nxt_u[PQ][c]);
      if (PUPDATE) settle(nxt_p[PQ][c]);
    }
    if (s < xc && active) {
      T un[3][ZPT];
      combine(nxt_u[PQ], nxt_p[PQ], un);
      if (PUPDATE && owns && s + 1 < xc) store_dir(plane_off(x0 + s + 1) + roff, un);
      x_stage(std::integral_constant<int, PQ>{}, un);
    }
    if (s + 2 < xc) advance(off_ld);
    if (s < xc - 1) advance(off_sv);
    advance(off_st);
    if (CHEB && s < xc) advance(off_xq);
    __syncthreads();
    return false;
  };

  for (int s = 0; s <= xc + 1; s += 2) {
    if (step(s, std::integral_constant<int, 0>{})) break;
    if (step(s + 1, std::integral_constant<int, 1>{})) break;
  }
#undef LOFF
  if (DOT) {
    const double t = block_sum(dot, red);
    if (threadIdx.x == 0) a.dot_part[(int64_t)rhs * a.nitems + a.item_base + blockIdx.x] = t;
  }
}

// Host-side launcher -------------------------------------------------------------------------

template <typename T, int ZPT, int DPP>
static hipError_t launch_mv_zpt(const MatvecArgs<T>& a, int strain, int pupdate, int dot, int cheb, dim3 grid,
                                int threads, size_t lds, hipStream_t st) {
  if (cheb) {
    if (!strain || pupdate || dot || a.s64 != nullptr) return hipErrorInvalidValue;
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)matvec_kernel<T, ZPT, true, false, false, DPP, false, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((matvec_kernel<T, ZPT, true, false, false, DPP, false, true>), grid, dim3(threads), lds, st, a);
    return hipGetLastError();
  }
#define NLFH_MV(S_, P_, D_)                                                                           \
  if (lds > 65536)                                                                                  \
    (void)hipFuncSetAttribute((const void*)matvec_kernel<T, ZPT, S_, P_, D_, DPP>,                     \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                \
  hipLaunchKernelGGL((matvec_kernel<T, ZPT, S_, P_, D_, DPP>), grid, dim3(threads), lds, st, a);       \
  return hipGetLastError();
  if (a.s64 != nullptr) {
    // the FP64-stiffness read serves the standalone FP32 operator only (no strain, p update or dot)
    if constexpr (sizeof(T) == 4) {
      if (!strain && !pupdate && !dot) {
        if (lds > 65536)
          (void)hipFuncSetAttribute((const void*)matvec_kernel<T, ZPT, false, false, false, DPP, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL((matvec_kernel<T, ZPT, false, false, false, DPP, true>), grid, dim3(threads), lds, st, a);
        return hipGetLastError();
      }
    }
    return hipErrorInvalidValue;
  }
  if (strain) {
    NLFH_MV(true, false, false)
  }
  if (pupdate) {
    if (dot) { NLFH_MV(false, true, true) }
    NLFH_MV(false, true, false)
  }
  if (dot) { NLFH_MV(false, false, true) }
  NLFH_MV(false, false, false)
#undef NLFH_MV
}

template <typename T>
hipError_t launch_matvec(const MatvecArgs<T>& a, int nblk, int zpt, int max_ty, int strain, int pupdate, int dot,
                         int cheb, hipStream_t st) {
  // max_ty + 2 rows of LZS lanes per block
  const int threads = ((a.LZS * (max_ty + 2) + 63) / 64) * 64;
  const size_t lds = (size_t)3 * (max_ty + 2) * 6 * zpt * a.LZ * sizeof(T);
  dim3 grid(nblk, a.m, 1);
  // DPP mode: 1 when a row is a whole wavefront, 2 when it is one 16-lane DPP row, else 0 (ds_bpermute)
  const int dpp = a.LZ == 64 ? 1 : (a.LZ == 16 && a.LZS == 16 ? 2 : 0);
  switch (zpt * 4 + dpp) {
    case 4: return launch_mv_zpt<T, 1, 0>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 5: return launch_mv_zpt<T, 1, 1>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 6: return launch_mv_zpt<T, 1, 2>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 8: return launch_mv_zpt<T, 2, 0>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 9: return launch_mv_zpt<T, 2, 1>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 10: return launch_mv_zpt<T, 2, 2>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 12: return launch_mv_zpt<T, 3, 0>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 14: return launch_mv_zpt<T, 3, 2>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 16: return launch_mv_zpt<T, 4, 0>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 17: return launch_mv_zpt<T, 4, 1>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    case 18: return launch_mv_zpt<T, 4, 2>(a, strain, pupdate, dot, cheb, grid, threads, lds, st);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_matvec<double>(const MatvecArgs<double>&, int, int, int, int, int, int, int, hipStream_t);
template hipError_t launch_matvec<float>(const MatvecArgs<float>&, int, int, int, int, int, int, int, hipStream_t);

}  // namespace nlfh
