// Periodic 3D convolution network kernels for the learned warm start (gfx950, BF16 MFMA with FP32
// accumulation). Activations are channels-last [batch][x][y][z][C] (z fastest), C a multiple of 16.
//
//   conv:   out[v][co] = bias[co] + sum_{tap, ci} W[tap][co][ci] in[v + d_tap][ci]     (periodic)
//           implicit GEMM on v_mfma_f32_16x16x16_bf16: a wavefront owns 16 voxels x all output
//           channels; K = 27 taps x input channels in steps of 16. Two input sources are
//           concatenated on the fly (decoder skip connections) and source A may be read at half
//           resolution (nearest upsampling fused into the load). flip = 1 reads d -> -d, which with
//           transposed weights is the data gradient of a convolution.
//   wgrad:  dW[tap][co][ci] = sum_v g[v][co] in[v + d_tap][ci]: per (tap, 16x16 block, voxel chunk)
//           one wavefront, partial sums per chunk, then a fixed-order reduction.
//   plus elementwise kernels (activation gradient, pooling, upsample-sum, loss, Adam, bias grads).
#include <cstdint>

#include "nlfh.h"
#include "nlfh_common.h"

namespace nlfh {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

struct ConvArgs {
  const uint16_t* a;  // source A [batch][na^3][ca]
  const uint16_t* b;  // source B [batch][n^3][cb] (nullable)
  const uint16_t* w;  // [27][coutp][cin], cin = ca + cb
  const float* bias;  // [cout] nullable
  float* out_f;       // [batch][n^3][cout] nullable
  uint16_t* out_h;    // [batch][n^3][cout] nullable (bf16 copy after the activation)
  int ca, cb, a_half;
  int n, batch, cout, coutp;
  int act;   // 0: none, 1: leaky relu (slope 0.1)
  int flip;  // 1: tap offsets negated
};

template <int NB>
__global__ void __launch_bounds__(256) conv3d_kernel(const ConvArgs p) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int n = p.n;
  const int64_t nv = (int64_t)n * n * n;
  const int64_t tile = ((int64_t)blockIdx.x * 4 + wave) * 16;  // first voxel (batch-flattened)
  if (tile >= nv * p.batch) return;
  const int bidx = (int)(tile / nv);
  const int64_t vb = tile - (int64_t)bidx * nv;
  const int row = lane & 15, kg = lane >> 4;
  // voxel of this lane in the A operand
  const int64_t va = vb + row;
  const int zx = (int)(va % n), yx = (int)((va / n) % n), xx = (int)(va / ((int64_t)n * n));
  const int cin = p.ca + p.cb;
  const int na = p.a_half ? n / 2 : n;
  f32x4 acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < 27; ++t) {
    int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    if (p.flip) { dx = -dx; dy = -dy; dz = -dz; }
    const int px = (xx + dx + n) % n, py = (yx + dy + n) % n, pz = (zx + dz + n) % n;
    const uint16_t* wt = p.w + (int64_t)t * p.coutp * cin;
    for (int c0 = 0; c0 < cin; c0 += 16) {
      const int c = c0 + 4 * kg;
      bf16x4 av;
      if (c0 < p.ca) {
        const int ax = p.a_half ? px / 2 : px, ay = p.a_half ? py / 2 : py, az = p.a_half ? pz / 2 : pz;
        const int64_t ai = (((int64_t)bidx * na + ax) * na + ay) * na + az;
        av = *reinterpret_cast<const bf16x4*>(p.a + ai * p.ca + c);
      } else {
        const int64_t bi = (((int64_t)bidx * n + px) * n + py) * n + pz;
        av = *reinterpret_cast<const bf16x4*>(p.b + bi * p.cb + (c - p.ca));
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const bf16x4 bv = *reinterpret_cast<const bf16x4*>(wt + (int64_t)(j * 16 + row) * cin + c);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[j], 0, 0, 0);
      }
    }
  }
  // D: lane holds rows (voxels) 4*kg + r, column (channel) j*16 + row
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int co = j * 16 + row;
    if (co >= p.cout) continue;
    const float bias = p.bias ? p.bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t v = tile + 4 * kg + r;
      float y = acc[j][r] + bias;
      if (p.act == 1) y = y > 0.f ? y : 0.1f * y;
      if (p.out_f) p.out_f[v * p.cout + co] = y;
      if (p.out_h) p.out_h[v * p.cout + co] = f2bf(y);
    }
  }
}

// LDS-staged form of the same convolution for cells whose side the brick tiles: a block owns a brick of
// 4 x BY x BZ = 256 output voxels (BZ = 16 when 16 | n, else 8 with BY = 8) and all output channels; per
// chunk of KC input channels (32 when both sources split into 32-channel chunks, else 16) it stages the
// brick with its one-voxel periodic halo in LDS (nearest upsampling of a half-resolution source applied
// as it is staged), then runs the 27 taps as K = KC steps of v_mfma_f32_16x16x32_bf16 (KC = 32) or
// v_mfma_f32_16x16x16_bf16 (KC = 16). A wavefront holds 4 M-tiles of 16 voxels (z-rows of the brick) x NB
// output-channel blocks; each weight fragment is loaded once per tap and used by the 4 M-tiles.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

template <int KC>
struct KFrag;
template <>
struct KFrag<32> {
  typedef bf16x8 T;
  __device__ static f32x4 mfma(T a, T b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct KFrag<16> {
  typedef bf16x4 T;
  __device__ static f32x4 mfma(T a, T b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
};

template <int BZ>
struct Brick {
  static constexpr int BX = 4, BY = 256 / (4 * BZ);
  static constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;
  static constexpr int HV = HX * HY * HZ;
  // voxel r (0..15) of M-tile mt (0..15) inside the brick
  __device__ static void voxel(int mt, int r, int& x, int& y, int& z) {
    x = mt >> 2;
    if (BZ == 16) {
      y = mt & 3;
      z = r;
    } else {
      y = (mt & 3) * 2 + (r >> 3);
      z = r & 7;
    }
  }
};

template <int KC, int NB, int BZ>
__global__ void __launch_bounds__(256) conv3d_lds_kernel(const ConvArgs p) {
  typedef Brick<BZ> Bk;
  typedef typename KFrag<KC>::T Frag;
  constexpr int KL = KC / 4;  // bf16 values per lane and fragment
  extern __shared__ uint4 lds_u4[];
  uint16_t* lds = reinterpret_cast<uint16_t*>(lds_u4);
  const int n = p.n;
  const int nbz = n / BZ, nby = n / Bk::BY, nbx = n / Bk::BX;
  int id = blockIdx.x;
  const int bz = id % nbz; id /= nbz;
  const int by = id % nby; id /= nby;
  const int bx = id % nbx;
  const int bidx = id / nbx;
  const int x0 = bx * Bk::BX, y0 = by * Bk::BY, z0 = bz * BZ;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int cin = p.ca + p.cb;
  const int na = p.a_half ? n / 2 : n;
  // LDS offset (in bf16) of the A-fragment voxel of each of this lane's M-tiles, tap (0, 0, 0)
  int abase[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    int x, y, z;
    Bk::voxel(wave * 4 + mt, row, x, y, z);
    abase[mt] = (((x + 1) * Bk::HY + (y + 1)) * Bk::HZ + (z + 1)) * KC + kg * KL;
  }
  f32x4 acc[4][NB];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[mt][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  constexpr int UPV = KC / 8;  // 16-byte units per staged voxel
  for (int c0 = 0; c0 < cin; c0 += KC) {
    const bool fromA = c0 < p.ca;
    for (int u = threadIdx.x; u < Bk::HV * UPV; u += 256) {
      const int hv = u / UPV, q = u - hv * UPV;
      const int hz = hv % Bk::HZ, hy = (hv / Bk::HZ) % Bk::HY, hx = hv / (Bk::HZ * Bk::HY);
      int X = x0 - 1 + hx, Y = y0 - 1 + hy, Z = z0 - 1 + hz;
      X = X < 0 ? X + n : (X >= n ? X - n : X);
      Y = Y < 0 ? Y + n : (Y >= n ? Y - n : Y);
      Z = Z < 0 ? Z + n : (Z >= n ? Z - n : Z);
      uint4 val;
      if (fromA) {
        const int ax = p.a_half ? X >> 1 : X, ay = p.a_half ? Y >> 1 : Y, az = p.a_half ? Z >> 1 : Z;
        const int64_t ai = (((int64_t)bidx * na + ax) * na + ay) * na + az;
        val = *reinterpret_cast<const uint4*>(p.a + ai * p.ca + c0 + q * 8);
      } else {
        const int64_t bi = (((int64_t)bidx * n + X) * n + Y) * n + Z;
        val = *reinterpret_cast<const uint4*>(p.b + bi * p.cb + (c0 - p.ca) + q * 8);
      }
      lds_u4[hv * UPV + q] = val;
    }
    __syncthreads();
    const uint16_t* wc = p.w + c0 + kg * KL;
    // The weight fragments rotate through three slots, three taps ahead: tap t's products read slot t mod 3, which
    // then takes tap t + 3's fragments, so a load has two taps' products to land; the loop runs the three taps of
    // one z-line per pass (27 = 9 x 3), so the slot is a compile-time index.
    Frag bw[3][NB];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < NB; ++j)
        bw[k][j] = *reinterpret_cast<const Frag*>(wc + ((int64_t)k * p.coutp + j * 16 + row) * cin);
#pragma unroll 1
    for (int t0 = 0; t0 < 27; t0 += 3) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int t = t0 + k;
        int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = k - 1;
        if (p.flip) { dx = -dx; dy = -dy; dz = -dz; }
        const int toff = ((dx * Bk::HY + dy) * Bk::HZ + dz) * KC;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          const Frag av = *reinterpret_cast<const Frag*>(lds + abase[mt] + toff);
#pragma unroll
          for (int j = 0; j < NB; ++j) acc[mt][j] = KFrag<KC>::mfma(av, bw[k][j], acc[mt][j]);
        }
        if (t + 3 < 27) {
#pragma unroll
          for (int j = 0; j < NB; ++j)
            bw[k][j] = *reinterpret_cast<const Frag*>(wc + ((int64_t)(t + 3) * p.coutp + j * 16 + row) * cin);
        }
      }
    }
    __syncthreads();
  }
  // D: lane holds voxels 4 kg + e of each M-tile, output channel j * 16 + row
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int co = j * 16 + row;
    if (co >= p.cout) continue;
    const float bias = p.bias ? p.bias[co] : 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int x, y, z;
        Bk::voxel(wave * 4 + mt, 4 * kg + e, x, y, z);
        const int64_t v = (((int64_t)bidx * n + x0 + x) * n + y0 + y) * n + z0 + z;
        float yv = acc[mt][j][e] + bias;
        if (p.act == 1) yv = yv > 0.f ? yv : 0.1f * yv;
        if (p.out_f) p.out_f[v * p.cout + co] = yv;
        if (p.out_h) p.out_h[v * p.cout + co] = f2bf(yv);
      }
  }
}

struct WgradArgs {
  const uint16_t* g;  // [batch][n^3][cg] output-side gradient (bf16)
  const uint16_t* a;  // source A [batch][na^3][ca]
  const uint16_t* b;  // source B [batch][n^3][cb]
  float* part;        // [nchunk][27][cg][cin]
  int ca, cb, a_half, cg;
  int n, batch;
  int64_t chunk;      // voxels per chunk (multiple of 16)
  int nchunk;
};

// one wavefront per (chunk, tap, co block, ci block): D[co][ci] += sum_k g[v_k][co] in[v_k + d][ci]
__global__ void __launch_bounds__(64) wgrad_kernel(const WgradArgs p) {
  const int lane = threadIdx.x;
  const int cin = p.ca + p.cb;
  const int nci = cin / 16, nco = p.cg / 16;
  int id = blockIdx.x;
  const int cib = id % nci; id /= nci;
  const int cob = id % nco; id /= nco;
  const int t = id % 27; id /= 27;
  const int ch = id;
  if (ch >= p.nchunk) return;
  const int n = p.n;
  const int64_t nv = (int64_t)n * n * n;
  const int64_t total = nv * p.batch;
  const int64_t v0 = (int64_t)ch * p.chunk;
  const int64_t v1 = v0 + p.chunk < total ? v0 + p.chunk : total;
  const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
  const int row = lane & 15, kg = lane >> 4;
  const int co = cob * 16 + row, ci = cib * 16 + row;
  const int na = p.a_half ? n / 2 : n;
  f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int64_t vk = v0; vk < v1; vk += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t v = vk + 4 * kg + j;
      uint16_t ga = 0, sb = 0;
      if (v < v1) {
        ga = p.g[v * p.cg + co];
        const int bidx = (int)(v / nv);
        const int64_t vv = v - (int64_t)bidx * nv;
        const int z = (int)(vv % n), y = (int)((vv / n) % n), x = (int)(vv / ((int64_t)n * n));
        const int px = (x + dx + n) % n, py = (y + dy + n) % n, pz = (z + dz + n) % n;
        if (ci < p.ca) {
          const int ax = p.a_half ? px / 2 : px, ay = p.a_half ? py / 2 : py, az = p.a_half ? pz / 2 : pz;
          sb = p.a[((((int64_t)bidx * na + ax) * na + ay) * na + az) * p.ca + ci];
        } else {
          sb = p.b[((((int64_t)bidx * n + px) * n + py) * n + pz) * p.cb + (ci - p.ca)];
        }
      }
      av[j] = (short)ga;
      bv[j] = (short)sb;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  // D rows = co (4*kg + r), columns = ci (row)
  float* out = p.part + (((int64_t)ch * 27 + t) * p.cg) * cin;
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(int64_t)(cob * 16 + 4 * kg + r) * cin + ci] = acc[r];
}

// Weight gradient over bricks of 4 x 4 x 16 voxels (cells with 16 | n): a block takes a run of bricks of one
// cell batch, one 16-channel input block (ci0) and one 16-channel output-gradient block (co0); per brick it
// stages g[v][co0..+16) (256 voxels) and the input's halo brick [6 x 6 x 18][ci0..+16) in LDS, and every
// wavefront accumulates D_t[co][ci] = sum_v g[v][co] in[v + d_t][ci] for its taps (t = wave, wave + 4, ...)
// with v_mfma_f32_16x16x16_bf16, K = the 16 voxels of one z-row of the brick. The g fragments of the 16 rows
// are read once per brick and reused by every tap. Partial sums per run: part[run][t][co][ci].
__global__ void __launch_bounds__(256) wgrad_lds_kernel(const WgradArgs p, int bricks_per_run) {
  typedef Brick<16> Bk;
  __shared__ uint4 sg[256 * 2];       // g brick: [voxel][16 channels]
  __shared__ uint4 sx[Bk::HV * 2];    // input halo brick: [halo voxel][16 channels]
  const uint16_t* g16 = reinterpret_cast<const uint16_t*>(sg);
  const uint16_t* x16 = reinterpret_cast<const uint16_t*>(sx);
  const int cin = p.ca + p.cb;
  const int nci = cin / 16, nco = p.cg / 16;
  int id = blockIdx.x;
  const int cib = id % nci; id /= nci;
  const int cob = id % nco; id /= nco;
  const int run = id;
  const int n = p.n;
  const int na = p.a_half ? n / 2 : n;
  const int nbz = n / 16, nby = n / 4, nbx = n / 4;
  const int64_t per_cell = (int64_t)nbz * nby * nbx;
  const int64_t nbr = per_cell * p.batch;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int ci0 = cib * 16, co0 = cob * 16;
  const bool fromA = ci0 < p.ca;
  f32x4 acc[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) acc[k] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int bi = 0; bi < bricks_per_run; ++bi) {
    const int64_t br = (int64_t)run * bricks_per_run + bi;
    if (br >= nbr) break;  // uniform over the block
    const int bidx = (int)(br / per_cell);
    int64_t r = br - (int64_t)bidx * per_cell;
    const int bz = (int)(r % nbz); r /= nbz;
    const int by = (int)(r % nby);
    const int bx = (int)(r / nby);
    const int x0 = bx * 4, y0 = by * 4, z0 = bz * 16;
    // stage: g rows (two 16-byte units per voxel), input halo (two units per halo voxel)
    for (int u = threadIdx.x; u < 512; u += 256) {
      const int v = u >> 1, q = u & 1;
      const int x = v >> 6, y = (v >> 4) & 3, z = v & 15;  // brick voxel order: M-tile (x, y), z
      const int64_t gv = (((int64_t)bidx * n + x0 + x) * n + y0 + y) * n + z0 + z;
      sg[u] = *reinterpret_cast<const uint4*>(p.g + gv * p.cg + co0 + q * 8);
    }
    for (int u = threadIdx.x; u < Bk::HV * 2; u += 256) {
      const int hv = u >> 1, q = u & 1;
      const int hz = hv % Bk::HZ, hy = (hv / Bk::HZ) % Bk::HY, hx = hv / (Bk::HZ * Bk::HY);
      int X = x0 - 1 + hx, Y = y0 - 1 + hy, Z = z0 - 1 + hz;
      X = X < 0 ? X + n : (X >= n ? X - n : X);
      Y = Y < 0 ? Y + n : (Y >= n ? Y - n : Y);
      Z = Z < 0 ? Z + n : (Z >= n ? Z - n : Z);
      uint4 val;
      if (fromA) {
        const int ax = p.a_half ? X >> 1 : X, ay = p.a_half ? Y >> 1 : Y, az = p.a_half ? Z >> 1 : Z;
        val = *reinterpret_cast<const uint4*>(p.a + ((((int64_t)bidx * na + ax) * na + ay) * na + az) * p.ca + ci0 +
                                              q * 8);
      } else {
        val = *reinterpret_cast<const uint4*>(p.b + ((((int64_t)bidx * n + X) * n + Y) * n + Z) * p.cb +
                                              (ci0 - p.ca) + q * 8);
      }
      sx[u] = val;
    }
    __syncthreads();
    // A fragments (M = output channel, K = 4 voxels of a z-row) for the 16 rows of the brick
    bf16x4 ga[16];
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
      for (int e = 0; e < 4; ++e) ga[m][e] = (short)g16[(m * 16 + kg * 4 + e) * 16 + row];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int t = wave + 4 * k;
      if (t < 27) {
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          const int x = m >> 2, y = m & 3;
          const int hb = ((x + 1 + dx) * Bk::HY + (y + 1 + dy)) * Bk::HZ + (1 + dz) + kg * 4;
          bf16x4 xb;
#pragma unroll
          for (int e = 0; e < 4; ++e) xb[e] = (short)x16[(hb + e) * 16 + row];
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ga[m], xb, acc[k], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // D rows = output channel 4 kg + e, column = input channel `row`
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const int t = wave + 4 * k;
    if (t >= 27) continue;
    float* out = p.part + (((int64_t)run * 27 + t) * p.cg) * cin;
#pragma unroll
    for (int e = 0; e < 4; ++e) out[(int64_t)(co0 + 4 * kg + e) * cin + ci0 + row] = acc[k][e];
  }
}

// Bias gradient, first level: block b sums voxels [b * chunk, (b + 1) * chunk) of dy * act'(y) per channel
// (threads = vl voxel lanes x C channels, coalesced over channels-last rows), fixed-order sum over the
// voxel lanes; part[b][c].
__global__ void bias_part_kernel(const float* dy, const uint16_t* y, int64_t nvox, int C, int act, int64_t chunk,
                                 float* part) {
  __shared__ float red[256];
  const int vl = blockDim.x / C;
  const int c = threadIdx.x % C, vi = threadIdx.x / C;
  const int64_t v0 = (int64_t)blockIdx.x * chunk, v1 = v0 + chunk < nvox ? v0 + chunk : nvox;
  float s = 0.f;
  for (int64_t v = v0 + vi; v < v1; v += vl) {
    float d = dy[v * C + c];
    if (act == 1 && bf2f(y[v * C + c]) <= 0.f) d *= 0.1f;
    s += d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (vi == 0) {
    float t = 0.f;
    for (int k = 0; k < vl; ++k) t += red[k * C + c];
    part[(int64_t)blockIdx.x * C + c] = t;
  }
}

// dW[i] = sum over chunks of part[chunk][i], in a fixed order: a block of 256 threads owns opb consecutive outputs
// (a power of two <= 64) and splits the chunks over its G = 256 / opb thread groups (group g sums chunks g, g + G,
// ... with eight loads in flight); the groups' sums are then added in group order. A thread per output summing
// every chunk in turn (up to 2048) was bound by the latency of its dependent loads: 175 us per call on average.
__global__ void __launch_bounds__(256) reduce_chunks_kernel(const float* part, int nchunk, int64_t len, float* out,
                                                            int opb) {
  __shared__ float red[256];
  const int G = 256 / opb;
  const int g = threadIdx.x / opb, o = threadIdx.x - g * opb;
  const int64_t i = (int64_t)blockIdx.x * opb + o;
  float s = 0.f;
  if (i < len) {
    int c = g;
    for (; c + 7 * G < nchunk; c += 8 * G) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = part[(int64_t)(c + k * G) * len + i];
#pragma unroll
      for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; c < nchunk; c += G) s += part[(int64_t)c * len + i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (g == 0 && i < len) {
    float t = 0.f;
    for (int k = 0; k < G; ++k) t += red[k * opb + o];
    out[i] = t;
  }
}

// outputs per block of reduce_chunks_kernel: the largest power of two <= 64 that still gives about 1024 blocks
static int reduce_opb(int64_t len) {
  int opb = 64;
  while (opb > 4 && (len + opb - 1) / opb < 1024) opb >>= 1;
  return opb;
}
static void launch_reduce_chunks(const float* part, int nchunk, int64_t len, float* out, hipStream_t st) {
  const int opb = reduce_opb(len);
  hipLaunchKernelGGL(reduce_chunks_kernel, dim3((unsigned)((len + opb - 1) / opb)), dim3(256), 0, st, part, nchunk, len,
                     out, opb);
}

// g = bf16(dy * act'(y)) with leaky relu slope 0.1 (act = 1) or identity (act = 0)
__global__ void act_grad_kernel(const float* dy, const uint16_t* y, uint16_t* g, int64_t len, int act) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    float d = dy[i];
    if (act == 1 && bf2f(y[i]) <= 0.f) d *= 0.1f;
    g[i] = f2bf(d);
  }
}

// bias gradient: db[c] = sum_v g[v][c] (g as f32 via dy * act'), one block per channel
__global__ void bias_grad_kernel(const float* dy, const uint16_t* y, int64_t nvox, int c_total, int act, float* db) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int64_t v = threadIdx.x; v < nvox; v += blockDim.x) {
    float d = dy[v * c_total + c];
    if (act == 1 && bf2f(y[v * c_total + c]) <= 0.f) d *= 0.1f;
    s += d;
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) db[c] = (float)t;
}

// 2x2x2 average pooling, bf16 -> bf16, [batch][n^3][C] -> [batch][(n/2)^3][C]
__global__ void pool_kernel(const uint16_t* in, uint16_t* out, int n, int batch, int C) {
  const int h = n / 2;
  const int64_t total = (int64_t)batch * h * h * h * C;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    int64_t v = i / C;
    const int z = (int)(v % h); v /= h;
    const int y = (int)(v % h); v /= h;
    const int x = (int)(v % h);
    const int b = (int)(v / h);
    float s = 0.f;
    for (int a = 0; a < 8; ++a) {
      const int xi = 2 * x + (a & 1), yi = 2 * y + ((a >> 1) & 1), zi = 2 * z + ((a >> 2) & 1);
      s += bf2f(in[((((int64_t)b * n + xi) * n + yi) * n + zi) * C + c]);
    }
    out[i] = f2bf(0.125f * s);
  }
}

// gradient of nearest upsampling / pooling: coarse[v][c] (+)= scale * sum of the 8 fine children
// (fine is [batch][n^3][cf] and channels [c0, c0 + C) are taken); add = 1 accumulates.
__global__ void upsum_kernel(const float* fine, int cf, int c0, float* coarse, int n, int batch, int C, float scale,
                             int add) {
  const int h = n / 2;
  const int64_t total = (int64_t)batch * h * h * h * C;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    int64_t v = i / C;
    const int z = (int)(v % h); v /= h;
    const int y = (int)(v % h); v /= h;
    const int x = (int)(v % h);
    const int b = (int)(v / h);
    float s = 0.f;
    for (int a = 0; a < 8; ++a) {
      const int xi = 2 * x + (a & 1), yi = 2 * y + ((a >> 1) & 1), zi = 2 * z + ((a >> 2) & 1);
      s += fine[((((int64_t)b * n + xi) * n + yi) * n + zi) * cf + c0 + c];
    }
    coarse[i] = (add ? coarse[i] : 0.f) + scale * s;
  }
}

// gradient of average pooling: fine[v][c] (+)= scale * coarse[v / 2][c]
__global__ void updist_kernel(const float* coarse, float* fine, int n, int batch, int C, float scale, int add) {
  const int h = n / 2;
  const int64_t total = (int64_t)batch * n * n * n * C;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    int64_t v = i / C;
    const int z = (int)(v % n); v /= n;
    const int y = (int)(v % n); v /= n;
    const int x = (int)(v % n);
    const int b = (int)(v / n);
    const float s = scale * coarse[((((int64_t)b * h + x / 2) * h + y / 2) * h + z / 2) * C + c];
    fine[i] = add ? fine[i] + s : s;
  }
}

// dst[v][c] (+)= src[v][c0 + c] for c < C (src has cs channels) -- channel slices of gradients
__global__ void slice_add_kernel(const float* src, int cs, int c0, float* dst, int64_t nvox, int C, int add) {
  const int64_t total = nvox * C;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const float s = src[v * cs + c0 + c];
    dst[i] = add ? dst[i] + s : s;
  }
}

// MSE loss on the first `cl` channels: loss partial per block and grad = 2 (pred - target) / count
__global__ void mse_kernel(const float* pred, const float* target, float* grad, int64_t len, double inv_count,
                           double* loss_part) {
  __shared__ double red[4];
  double s = 0.0;
  const float gscale = (float)(2.0 * inv_count);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t done = 0;
  const bool vec = ((reinterpret_cast<uintptr_t>(pred) | reinterpret_cast<uintptr_t>(target) |
                     reinterpret_cast<uintptr_t>(grad)) & 15) == 0;
  if (vec) {
    // 16-byte vectors, four per thread in flight (a scalar grid-stride loop of one load pair per step was bound
    // by load latency: 253 us for 4 cells of 64^3 x 18 channels)
    const float4* P = reinterpret_cast<const float4*>(pred);
    const float4* T = reinterpret_cast<const float4*>(target);
    float4* G = reinterpret_cast<float4*>(grad);
    const int64_t n4 = len / 4;
    int64_t i = tid;
    for (; i + 3 * stride < n4; i += 4 * stride) {
      float4 p[4], t[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        p[k] = P[i + k * stride];
        t[k] = T[i + k * stride];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float4 d = {p[k].x - t[k].x, p[k].y - t[k].y, p[k].z - t[k].z, p[k].w - t[k].w};
        s += (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z + (double)d.w * d.w;
        G[i + k * stride] = {gscale * d.x, gscale * d.y, gscale * d.z, gscale * d.w};
      }
    }
    for (; i < n4; i += stride) {
      const float4 p = P[i], t = T[i];
      const float4 d = {p.x - t.x, p.y - t.y, p.z - t.z, p.w - t.w};
      s += (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z + (double)d.w * d.w;
      G[i] = {gscale * d.x, gscale * d.y, gscale * d.z, gscale * d.w};
    }
    done = n4 * 4;
  }
  for (int64_t i = done + tid; i < len; i += stride) {
    const float d = pred[i] - target[i];
    s += (double)d * d;
    grad[i] = gscale * d;
  }
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) loss_part[blockIdx.x] = t * inv_count;
}

// Network input: x[v][0] = bf16(2 rho[v] - 1) (rho rounded to float32 first, as the float32 path does), channels
// 1..15 zero; two 16-byte stores per voxel.
__global__ void input_kernel(const double* rho, int64_t nvox, uint4* x) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nvox; v += (int64_t)gridDim.x * blockDim.x) {
    const float r = (float)rho[v];
    const uint16_t h = f2bf(2.0f * r - 1.0f);
    x[2 * v] = make_uint4((uint32_t)h, 0u, 0u, 0u);
    x[2 * v + 1] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// Read-out of the network head for the learned start: dst[b][c][v] = scale * out[b][v][c] - mean_v(scale * out[b][v][c])
// in float64 for the channels c < nch of the ldc-wide float32 head, channel-major per cell (the planar field layout).
// Three launches with fixed summation orders: per (cell, chunk of voxels) channel sums (8 voxel groups x 32 channel
// lanes per block, four loads in flight, groups combined in order), per-cell means from the chunk sums in chunk
// order, and the scaled, mean-free write (a thread per voxel).
constexpr int kReadoutChunks = 512;
__global__ void __launch_bounds__(256) readout_part_kernel(const float* out, int ldc, int nch, int64_t nvox,
                                                           double* part) {
  __shared__ double red[256];
  const int b = blockIdx.y, chunk = blockIdx.x;
  const int g = threadIdx.x >> 5, c = threadIdx.x & 31;
  const int64_t len = (nvox + kReadoutChunks - 1) / kReadoutChunks;
  const int64_t v0 = chunk * len, v1 = v0 + len < nvox ? v0 + len : nvox;
  const float* o = out + (int64_t)b * nvox * ldc + c;
  double s = 0.0;
  if (c < nch) {
    int64_t v = v0 + g;
    for (; v + 24 < v1; v += 32) {
      float x[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) x[k] = o[(v + 8 * k) * ldc];
#pragma unroll
      for (int k = 0; k < 4; ++k) s += (double)x[k];
    }
    for (; v < v1; v += 8) s += (double)o[v * ldc];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (g == 0 && c < nch) {
    double t = 0.0;
    for (int k = 0; k < 8; ++k) t += red[k * 32 + c];
    part[((int64_t)b * kReadoutChunks + chunk) * nch + c] = t;
  }
}

__global__ void __launch_bounds__(256) readout_mean_kernel(const double* part, int nch, int64_t nvox, double scale,
                                                           double* mean) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const int g = threadIdx.x >> 5, c = threadIdx.x & 31;
  double s = 0.0;
  if (c < nch)
    for (int k = g; k < kReadoutChunks; k += 8) s += part[((int64_t)b * kReadoutChunks + k) * nch + c];
  red[threadIdx.x] = s;
  __syncthreads();
  if (g == 0 && c < nch) {
    double t = 0.0;
    for (int k = 0; k < 8; ++k) t += red[k * 32 + c];
    mean[(int64_t)b * nch + c] = scale * t / (double)nvox;
  }
}

__global__ void __launch_bounds__(256) readout_write_kernel(const float* out, int ldc, int nch, int64_t nvox,
                                                            double scale, const double* mean, double* dst) {
  const int b = blockIdx.y;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  const float* o = out + ((int64_t)b * nvox + v) * ldc;
  double* d = dst + (int64_t)b * nch * nvox + v;
  const double* mb = mean + (int64_t)b * nch;
  for (int c = 0; c < nch; ++c) d[(int64_t)c * nvox] = (double)o[c] * scale - mb[c];
}

// Adam step on FP32 master weights; refreshes the BF16 copy. Weights laid out [27][coutp][cin];
// entries with co >= cout stay zero (padding of the output channels).
__global__ void adam_kernel(float* w, const float* g, float* m1, float* m2, uint16_t* wh, int64_t len, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * w[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float upd = lr * (a / c1) / (sqrtf(b / c2) + eps);
    const float wn = w[i] - upd;
    w[i] = wn;
    if (wh) wh[i] = f2bf(wn);
  }
}

__global__ void f2bf_kernel(const float* in, uint16_t* out, int64_t len) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = f2bf(in[i]);
}

// transpose-and-flip of weights for the data gradient: wt[26 - t][ci][co] = w[t][co][ci]
// (output [27][cinp][coutp_src] with zero padding for ci >= cin)
__global__ void wflip_kernel(const uint16_t* w, int coutp, int cin, uint16_t* wt, int cinp) {
  const int64_t total = (int64_t)27 * cinp * coutp;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int co = (int)(i % coutp);
    const int64_t r = i / coutp;
    const int ci = (int)(r % cinp);
    const int t = (int)(r / cinp);
    wt[i] = ci < cin ? w[((int64_t)(26 - t) * coutp + co) * cin + ci] : (uint16_t)0;
  }
}

static inline int nb_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace nlfh

using namespace nlfh;

extern "C" {

int nlfh_net_conv(const uint16_t* a, int ca, int a_half, const uint16_t* b, int cb, const uint16_t* w,
                  const float* bias, int n, int batch, int cout, int coutp, int act, int flip, float* out_f,
                  uint16_t* out_h, void* stream) {
  if (!a || !w || (cb && !b) || ca % 16 || cb % 16 || coutp % 16 || coutp < cout || coutp > 96 || n < 2 ||
      (a_half && n % 2) || batch < 1 || (n * n * n) % 16) {
    set_error("nlfh_net_conv: invalid arguments (channels must be multiples of 16, coutp <= 96)");
    return NLFH_EINVAL;
  }
  ConvArgs p{a, b, w, bias, out_f, out_h, ca, cb, a_half, n, batch, cout, coutp, act, flip};
  hipStream_t st = (hipStream_t)stream;
  const int bz = n % 16 == 0 ? 16 : (n % 8 == 0 ? 8 : 0);
  if (bz) {  // the LDS-staged brick kernel
    const int kc = (ca % 32 == 0 && cb % 32 == 0) ? 32 : 16;
    const int64_t nblk = (int64_t)batch * (n / 4) * (n / (256 / (4 * bz))) * (n / bz);
    const size_t lds = (size_t)(4 + 2) * (256 / (4 * bz) + 2) * (bz + 2) * kc * sizeof(uint16_t);
    const int nb = coutp / 16;
#define NLFH_CONV_LDS(KC_, BZ_)                                                                            \
  switch (nb) {                                                                                          \
    case 1: hipLaunchKernelGGL((conv3d_lds_kernel<KC_, 1, BZ_>), dim3((unsigned)nblk), dim3(256), lds, st, p); break; \
    case 2: hipLaunchKernelGGL((conv3d_lds_kernel<KC_, 2, BZ_>), dim3((unsigned)nblk), dim3(256), lds, st, p); break; \
    case 3: hipLaunchKernelGGL((conv3d_lds_kernel<KC_, 3, BZ_>), dim3((unsigned)nblk), dim3(256), lds, st, p); break; \
    case 4: hipLaunchKernelGGL((conv3d_lds_kernel<KC_, 4, BZ_>), dim3((unsigned)nblk), dim3(256), lds, st, p); break; \
    case 5: hipLaunchKernelGGL((conv3d_lds_kernel<KC_, 5, BZ_>), dim3((unsigned)nblk), dim3(256), lds, st, p); break; \
    default: hipLaunchKernelGGL((conv3d_lds_kernel<KC_, 6, BZ_>), dim3((unsigned)nblk), dim3(256), lds, st, p); break; \
  }
    if (kc == 32 && bz == 16) { NLFH_CONV_LDS(32, 16) }
    else if (kc == 32) { NLFH_CONV_LDS(32, 8) }
    else if (bz == 16) { NLFH_CONV_LDS(16, 16) }
    else { NLFH_CONV_LDS(16, 8) }
#undef NLFH_CONV_LDS
    HIP_TRY(hipGetLastError());
    return NLFH_OK;
  }
  const int64_t tiles = ((int64_t)n * n * n * batch + 15) / 16;
  const int blocks = (int)((tiles + 3) / 4);
  switch (coutp / 16) {
    case 1: hipLaunchKernelGGL(conv3d_kernel<1>, dim3(blocks), dim3(256), 0, st, p); break;
    case 2: hipLaunchKernelGGL(conv3d_kernel<2>, dim3(blocks), dim3(256), 0, st, p); break;
    case 3: hipLaunchKernelGGL(conv3d_kernel<3>, dim3(blocks), dim3(256), 0, st, p); break;
    case 4: hipLaunchKernelGGL(conv3d_kernel<4>, dim3(blocks), dim3(256), 0, st, p); break;
    case 5: hipLaunchKernelGGL(conv3d_kernel<5>, dim3(blocks), dim3(256), 0, st, p); break;
    default: hipLaunchKernelGGL(conv3d_kernel<6>, dim3(blocks), dim3(256), 0, st, p); break;
  }
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_wgrad(const uint16_t* g, int cg, const uint16_t* a, int ca, int a_half, const uint16_t* b, int cb,
                   int n, int batch, float* part, long long part_len, float* dw, void* stream) {
  const int cin = ca + cb;
  if (!g || !a || (cb && !b) || cg % 16 || ca % 16 || cb % 16 || (n * n * n) % 16) {
    set_error("nlfh_net_wgrad: invalid arguments");
    return NLFH_EINVAL;
  }
  const int64_t total = (int64_t)n * n * n * batch;
  const int64_t wlen = (int64_t)27 * cg * cin;
  int64_t chunk = 4096;
  int64_t nchunk = (total + chunk - 1) / chunk;
  while (nchunk * wlen > part_len && chunk < total) {
    chunk *= 2;
    nchunk = (total + chunk - 1) / chunk;
  }
  if (nchunk * wlen > part_len) {
    set_error("nlfh_net_wgrad: partial buffer too small");
    return NLFH_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  if (n % 16 == 0) {  // bricks of 4 x 4 x 16 voxels, runs of bricks per block
    const int64_t nbr = (int64_t)batch * (n / 16) * (n / 4) * (n / 4);
    const int64_t pairs = (int64_t)(cin / 16) * (cg / 16);
    int64_t per_run = 1;
    while ((nbr + per_run - 1) / per_run * pairs > 2048 && per_run < nbr) per_run *= 2;  // about 2048 blocks
    int64_t nrun = (nbr + per_run - 1) / per_run;
    while (nrun * wlen > part_len && per_run < nbr) {
      per_run *= 2;
      nrun = (nbr + per_run - 1) / per_run;
    }
    if (nrun * wlen > part_len) {
      set_error("nlfh_net_wgrad: partial buffer too small");
      return NLFH_EINVAL;
    }
    WgradArgs p{g, a, b, part, ca, cb, a_half, cg, n, batch, 0, (int)nrun};
    hipLaunchKernelGGL(wgrad_lds_kernel, dim3((unsigned)(nrun * pairs)), dim3(256), 0, st, p, (int)per_run);
    launch_reduce_chunks(part, (int)nrun, wlen, dw, st);
    HIP_TRY(hipGetLastError());
    return NLFH_OK;
  }
  WgradArgs p{g, a, b, part, ca, cb, a_half, cg, n, batch, chunk, (int)nchunk};
  const int64_t blocks = nchunk * 27 * (cg / 16) * (cin / 16);
  hipLaunchKernelGGL(wgrad_kernel, dim3((unsigned)blocks), dim3(64), 0, st, p);
  launch_reduce_chunks(part, (int)nchunk, wlen, dw, st);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_act_grad(const float* dy, const uint16_t* y, uint16_t* g, long long len, int act, void* stream) {
  hipLaunchKernelGGL(act_grad_kernel, dim3(nb_for(len)), dim3(256), 0, (hipStream_t)stream, dy, y, g, (int64_t)len, act);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_bias_grad(const float* dy, const uint16_t* y, long long nvox, int c, int act, float* part,
                       long long part_len, float* db, void* stream) {
  if (!dy || !db || (act && !y) || c < 1 || c > 256 || nvox < 1) {
    set_error("nlfh_net_bias_grad: invalid arguments");
    return NLFH_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  if (part == nullptr || part_len < (long long)c) {  // no scratch: one block per channel
    hipLaunchKernelGGL(bias_grad_kernel, dim3(c), dim3(256), 0, st, dy, y, (int64_t)nvox, c, act, db);
    HIP_TRY(hipGetLastError());
    return NLFH_OK;
  }
  const int vl = 256 / c > 0 ? 256 / c : 1;
  int64_t nblk = (nvox + 511) / 512;  // about 512 voxels per block: enough blocks to cover the load latency
  if (nblk * c > part_len) nblk = part_len / c;
  const int64_t chunk = (nvox + nblk - 1) / nblk;
  nblk = (nvox + chunk - 1) / chunk;
  hipLaunchKernelGGL(bias_part_kernel, dim3((unsigned)nblk), dim3(vl * c), 0, st, dy, y, (int64_t)nvox, c, act, chunk,
                     part);
  launch_reduce_chunks(part, (int)nblk, (int64_t)c, db, st);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_pool(const uint16_t* in, uint16_t* out, int n, int batch, int c, void* stream) {
  const int64_t total = (int64_t)batch * (n / 2) * (n / 2) * (n / 2) * c;
  hipLaunchKernelGGL(pool_kernel, dim3(nb_for(total)), dim3(256), 0, (hipStream_t)stream, in, out, n, batch, c);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_upsum(const float* fine, int cf, int c0, float* coarse, int n, int batch, int c, float scale, int add,
                   void* stream) {
  const int64_t total = (int64_t)batch * (n / 2) * (n / 2) * (n / 2) * c;
  hipLaunchKernelGGL(upsum_kernel, dim3(nb_for(total)), dim3(256), 0, (hipStream_t)stream, fine, cf, c0, coarse, n,
                     batch, c, scale, add);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_updist(const float* coarse, float* fine, int n, int batch, int c, float scale, int add, void* stream) {
  const int64_t total = (int64_t)batch * n * n * n * c;
  hipLaunchKernelGGL(updist_kernel, dim3(nb_for(total)), dim3(256), 0, (hipStream_t)stream, coarse, fine, n, batch, c,
                     scale, add);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_slice_add(const float* src, int cs, int c0, float* dst, long long nvox, int c, int add, void* stream) {
  hipLaunchKernelGGL(slice_add_kernel, dim3(nb_for(nvox * c)), dim3(256), 0, (hipStream_t)stream, src, cs, c0, dst,
                     (int64_t)nvox, c, add);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_mse(const float* pred, const float* target, float* grad, long long len, double* loss_part, int nparts,
                 void* stream) {
  hipLaunchKernelGGL(mse_kernel, dim3(nparts), dim3(256), 0, (hipStream_t)stream, pred, target, grad, (int64_t)len,
                     1.0 / (double)len, loss_part);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_input(const double* rho, long long nvox, unsigned short* x, void* stream) {
  if (!rho || !x || nvox < 1 || (reinterpret_cast<uintptr_t>(x) & 15)) {
    set_error("nlfh_net_input: invalid arguments (x 16-byte aligned, nvox >= 1)");
    return NLFH_EINVAL;
  }
  hipLaunchKernelGGL(input_kernel, dim3(nb_for(nvox)), dim3(256), 0, (hipStream_t)stream, rho, (int64_t)nvox,
                     reinterpret_cast<uint4*>(x));
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_readout(const float* out, int ldc, int nch, long long nvox, int batch, double scale, double* part,
                     long long part_len, double* dst, void* stream) {
  if (!out || !part || !dst || nch < 1 || nch > 32 || ldc < nch || nvox < 1 || batch < 1 || batch > 65535 ||
      part_len < (long long)batch * (kReadoutChunks + 1) * nch) {
    set_error("nlfh_net_readout: invalid arguments (1 <= nch <= 32, ldc >= nch, part_len >= batch * 513 * nch)");
    return NLFH_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  double* mean = part + (int64_t)batch * kReadoutChunks * nch;
  hipLaunchKernelGGL(readout_part_kernel, dim3(kReadoutChunks, batch), dim3(256), 0, st, out, ldc, nch,
                     (int64_t)nvox, part);
  hipLaunchKernelGGL(readout_mean_kernel, dim3(batch), dim3(256), 0, st, part, nch, (int64_t)nvox, scale, mean);
  hipLaunchKernelGGL(readout_write_kernel, dim3((unsigned)((nvox + 255) / 256), batch), dim3(256), 0, st, out, ldc,
                     nch, (int64_t)nvox, scale, mean, dst);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_adam(float* w, const float* g, float* m1, float* m2, uint16_t* wh, long long len, float lr, float b1,
                  float b2, float eps, int step, float wd, void* stream) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(nb_for(len)), dim3(256), 0, (hipStream_t)stream, w, g, m1, m2, wh,
                     (int64_t)len, lr, b1, b2, eps, c1, c2, wd);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_f2bf(const float* in, uint16_t* out, long long len, void* stream) {
  hipLaunchKernelGGL(f2bf_kernel, dim3(nb_for(len)), dim3(256), 0, (hipStream_t)stream, in, out, (int64_t)len);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_net_wflip(const uint16_t* w, int coutp, int cin, uint16_t* wt, int cinp, void* stream) {
  const int64_t total = (int64_t)27 * cinp * coutp;
  hipLaunchKernelGGL(wflip_kernel, dim3(nb_for(total)), dim3(256), 0, (hipStream_t)stream, w, coutp, cin, wt, cinp);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

}  // extern "C"
