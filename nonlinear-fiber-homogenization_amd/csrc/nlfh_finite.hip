// Finite-strain cell operators on voxel grids (FP64): the cell energy, its gradient in the periodic
// displacement w and in the mean deformation gradient Fbar, the Hessian-vector product with the
// tangent at a state, applied matrix-free, and the per-element density derivative of the stress work
// <P, dF> that the adjoint density gradient needs (Mixed.grad_rho, mech.py:348-363).
//
// Follows nlfiber/mech.py: Q1 hexahedra with 2x2x2 Gauss points (mech.py:124-132, the element
// gradients of linear.py:33-42), the energy W = wq sum_e s_e sum_g psi_eff(Fbar + H_g) with wq = 1/(8N)
// (mech.py:141-154), the Wang-Lazarov-Sigmund-Jensen void interpolation psi_eff (mech.py:134-139,
// 107-112), the neo-Hookean and linear materials (mech.py:39-50) and the stored-tangent product of
// Mixed.hvp_A (mech.py:227-233; tangent_g mech.py:156-167).
//
// The reference stores the 3x3x3x3 tangent of every Gauss point (81 doubles x 8 points per voxel) and
// contracts it in each Hessian-vector product; here the product recomputes the state's deformation
// gradient from w and applies the tangent in closed form,
//   neo-Hookean:  dP = mu dF + lam tr(F^-1 dF) F^-T + (mu - lam ln J) (F^-1 dF F^-1)^T
//   linear:       dP = lam tr(dF) I + mu (dF + dF^T),
// with the void interpolation dP = gamma^2 (dP_nl(Fg) - dP_lin) + dP_lin, Fg = I + gamma (F - I), so
// each product reads two displacement fields instead of 5 KiB of tangent per voxel.
//
// Layout: w and v are the reference's (nx, ny, nz, 3) C-order arrays (component fastest); element e =
// voxel (i, j, k), e = (i ny + j) nz + k, spans nodes (i + a, j + b, k + c) mod n, node index
// a + 2b + 4c. Eight lanes per element (one per Gauss point) write the element's 24 nodal forces to a
// (24, N) buffer; a second kernel gathers the 8 contributions of every node in a fixed order, so results
// are deterministic.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "nlfh_common.h"

namespace nlfh {
namespace {

constexpr int kFsThreads = 256;
constexpr int kFsParts = 10;  // per-block partial sums: energy (or 0), then the 9 entries of sum P (or sum dP)

constexpr double kGaussOff = 0.28867513459481288225;  // 1 / (2 sqrt 3)

__host__ __device__ constexpr double gauss_coord(int bit) { return bit ? 0.5 + kGaussOff : 0.5 - kGaussOff; }

// d N_a / d x_d at Gauss point g of the unit element (linear.py:44-53): g and a count x fastest
__host__ __device__ constexpr double dshape(int g, int a, int d) {
  const double x = gauss_coord(g & 1), y = gauss_coord((g >> 1) & 1), z = gauss_coord((g >> 2) & 1);
  const int i = a & 1, j = (a >> 1) & 1, k = (a >> 2) & 1;
  const double fx = i ? x : 1.0 - x, fy = j ? y : 1.0 - y, fz = k ? z : 1.0 - z;
  const double sx = i ? 1.0 : -1.0, sy = j ? 1.0 : -1.0, sz = k ? 1.0 : -1.0;
  return d == 0 ? sx * fy * fz : (d == 1 ? fx * sy * fz : fx * fy * sz);
}

struct FsMat {
  double lam, mu;
  double emin, p;     // s = emin + (1 - emin) clip01(rho)^p
  double wls_b, wls_r0;
  int neo;            // 1 neo-Hookean, 0 linear
  int wls;            // void interpolation on
};

__device__ __forceinline__ double clip01(double x) { return x > 1.0 ? 1.0 : (x < 0.0 ? 0.0 : x); }

// psi_linear(F) and its first Piola stress (mech.py:39-43)
__device__ __forceinline__ double lin_psi_P(const double (&F)[3][3], double lam, double mu, double (&P)[3][3]) {
  double e[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) e[i][j] = 0.5 * ((F[i][j] - (i == j)) + (F[j][i] - (j == i)));
  const double tr = e[0][0] + e[1][1] + e[2][2];
  double ee = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ee += e[i][j] * e[i][j];
      P[i][j] = 2.0 * mu * e[i][j] + (i == j ? lam * tr : 0.0);
    }
  return 0.5 * lam * tr * tr + mu * ee;
}

__device__ __forceinline__ void lin_dP(const double (&dF)[3][3], double lam, double mu, double (&dP)[3][3]) {
  const double tr = dF[0][0] + dF[1][1] + dF[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) dP[i][j] = mu * (dF[i][j] + dF[j][i]) + (i == j ? lam * tr : 0.0);
}

// J = det F and F^-1 (adjugate / J)
__device__ __forceinline__ double inv3(const double (&F)[3][3], double (&Fi)[3][3]) {
  const double c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
  const double c01 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
  const double c02 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
  const double J = F[0][0] * c00 + F[0][1] * c01 + F[0][2] * c02;
  const double r = 1.0 / J;
  Fi[0][0] = c00 * r;
  Fi[1][0] = c01 * r;
  Fi[2][0] = c02 * r;
  Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * r;
  Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * r;
  Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * r;
  Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * r;
  Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * r;
  Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * r;
  return J;
}

// psi_neohooke(F) (mech.py:46-50) and P = mu F + (lam ln J - mu) F^-T; returns psi, fills Fi and lnJ
__device__ __forceinline__ double neo_psi_P(const double (&F)[3][3], double lam, double mu, double (&P)[3][3],
                                            double (&Fi)[3][3], double& lnJ) {
  const double J = inv3(F, Fi);
  lnJ = log(J);  // NaN for J <= 0, as in the reference (the line search rejects non-finite energies)
  double I1 = 0.0;
  const double c = lam * lnJ - mu;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      I1 += F[i][j] * F[i][j];
      P[i][j] = mu * F[i][j] + c * Fi[j][i];
    }
  return 0.5 * mu * (I1 - 3.0) - mu * lnJ + 0.5 * lam * lnJ * lnJ;
}

// tangent of the neo-Hookean energy at F (given F^-1, ln J) applied to dF
__device__ __forceinline__ void neo_dP(const double (&Fi)[3][3], double lnJ, const double (&dF)[3][3], double lam,
                                       double mu, double (&dP)[3][3]) {
  double A[3][3];  // F^-1 dF
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = Fi[i][0] * dF[0][j] + Fi[i][1] * dF[1][j] + Fi[i][2] * dF[2][j];
  const double tr = A[0][0] + A[1][1] + A[2][2];
  const double c = mu - lam * lnJ;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double B = A[j][0] * Fi[0][i] + A[j][1] * Fi[1][i] + A[j][2] * Fi[2][i];  // (F^-1 dF F^-1)_ji
      dP[i][j] = mu * dF[i][j] + lam * tr * Fi[j][i] + c * B;
    }
}

// psi_eff and its stress at F for one element's material (mech.py:134-139)
__device__ __forceinline__ double eff_psi_P(const double (&F)[3][3], const FsMat& m, double gam, double (&P)[3][3]) {
  if (!m.wls) {
    if (m.neo) {
      double Fi[3][3], lnJ;
      return neo_psi_P(F, m.lam, m.mu, P, Fi, lnJ);
    }
    return lin_psi_P(F, m.lam, m.mu, P);
  }
  double Fg[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Fg[i][j] = (i == j) + gam * (F[i][j] - (i == j));
  double Pn[3][3], Pg[3][3], Pl[3][3];
  double psi_n;
  if (m.neo) {
    double Fi[3][3], lnJ;
    psi_n = neo_psi_P(Fg, m.lam, m.mu, Pn, Fi, lnJ);
  } else {
    psi_n = lin_psi_P(Fg, m.lam, m.mu, Pn);
  }
  const double psi_g = lin_psi_P(Fg, m.lam, m.mu, Pg);
  const double psi_l = lin_psi_P(F, m.lam, m.mu, Pl);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = gam * (Pn[i][j] - Pg[i][j]) + Pl[i][j];
  return psi_n - psi_g + psi_l;
}

// tangent of psi_eff at F applied to dF
__device__ __forceinline__ void eff_dP(const double (&F)[3][3], const FsMat& m, double gam, const double (&dF)[3][3],
                                       double (&dP)[3][3]) {
  double Dl[3][3];
  lin_dP(dF, m.lam, m.mu, Dl);
  if (!m.neo) {  // psi_nl = psi_lin: the interpolation cancels
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dP[i][j] = Dl[i][j];
    return;
  }
  const double g = m.wls ? gam : 1.0;
  double Fg[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Fg[i][j] = m.wls ? (i == j) + g * (F[i][j] - (i == j)) : F[i][j];
  double Fi[3][3];
  const double lnJ = log(inv3(Fg, Fi));
  double Dn[3][3];
  neo_dP(Fi, lnJ, dF, m.lam, m.mu, Dn);
  if (!m.wls) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dP[i][j] = Dn[i][j];
    return;
  }
  const double g2 = g * g;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) dP[i][j] = g2 * (Dn[i][j] - Dl[i][j]) + Dl[i][j];
}

// P_eff at F and its derivative in the interpolation factor gamma (mech.py:134-139):
// dP/dgamma = (P_nl(Fg) - P_lin(Fg)) + gamma (A_nl(Fg) - A_lin) : (F - I), Fg = I + gamma (F - I); zero for the
// linear material (the interpolation cancels) and without the interpolation (gamma is not a function of rho)
__device__ __forceinline__ void eff_P_dgamma(const double (&F)[3][3], const FsMat& m, double gam, double (&P)[3][3],
                                             double (&Pg)[3][3]) {
  (void)eff_psi_P(F, m, gam, P);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Pg[i][j] = 0.0;
  if (!m.wls || !m.neo) return;
  double H[3][3], Fg[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      H[i][j] = F[i][j] - (i == j);
      Fg[i][j] = (i == j) + gam * H[i][j];
    }
  double Pn[3][3], Pl[3][3], Fi[3][3], lnJ;
  (void)neo_psi_P(Fg, m.lam, m.mu, Pn, Fi, lnJ);
  (void)lin_psi_P(Fg, m.lam, m.mu, Pl);
  double Dn[3][3], Dl[3][3];
  neo_dP(Fi, lnJ, H, m.lam, m.mu, Dn);
  lin_dP(H, m.lam, m.mu, Dl);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Pg[i][j] = (Pn[i][j] - Pl[i][j]) + gam * (Dn[i][j] - Dl[i][j]);
}

// Division by a divisor fixed at launch, n / d for any 32-bit n, by a multiply-high and shifts (the round-up
// method of Granlund and Montgomery, "Division by invariant integers using multiplication", PLDI 1994, sec. 4):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n), q = (t + ((n - t) >> s1)) >> s2 with
// s1 = min(l, 1), s2 = max(l - 1, 0) (d = 1 gives m = 1, t = 0, q = n).
struct FastDiv {
  unsigned d = 1, m = 1;
  int s1 = 0, s2 = 0;
  FastDiv() = default;
  explicit FastDiv(unsigned dv) : d(dv) {
    int l = 0;
    while ((1ull << l) < dv) ++l;
    m = (unsigned)((((1ull << l) - dv) << 32) / dv + 1);
    s1 = l < 1 ? l : 1;
    s2 = l > 1 ? l - 1 : 0;
  }
  __host__ __device__ unsigned div(unsigned n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned t = __umulhi(m, n);
#else
    const unsigned t = (unsigned)(((unsigned long long)m * n) >> 32);
#endif
    return (t + ((n - t) >> s1)) >> s2;
  }
};

struct FsArgs {
  int nx, ny, nz;
  int64_t n;
  FastDiv dz, dy;    // division by nz and by ny (element index -> voxel coordinates)
  const double* sg;  // (4, N): wq s_e, gamma_e, wq ds_e/drho_e, dgamma_e/drho_e of every element (fs_prepare_kernel)
  FsMat m;
  double wq;
  const double* w;   // state (N, 3)
  const double* v;   // direction (N, 3) of the product, or null (zero)
  double Fbar[9];
  const double* dFbar;  // device, 9 values (mode 2)
  double* fe;        // (24, N) element nodal forces, or null
  double* part;      // (kFsParts, blocks) partial sums
  unsigned pmask;    // partial sums to form (bit q: scalar q); the others are zero
};

// Eight consecutive lanes per element, lane g at Gauss point g. Lane g handles node c under the label
// j = c ^ g: the shape gradient of node c at point g is then sigma_g (.) T_j, where T_j depends on j only
// (the weight of a node along an axis is 1/2 + 1/(2 sqrt 3) when node and point sit on the same side, whatever
// the side) and sigma_g flips the sign of axis d when bit d of g is set. The element's 24 nodal forces
// sum_g P_g grad N_c(g) come out of a reduce-scatter over the 8 lanes in which every lane keeps labels
// j < 4 (then j < 2, then j = 0) and adds the partner's j + 4 (j + 2, j + 1): the partner's label for the
// same node (xor 4, 2, 1 of the lane) is exactly that. Lane g ends with label 0, node g. The order of the
// additions is fixed, so results are deterministic.
// MODE 0: energy; 1: energy, element forces of the gradient, sum of P; 2: Hessian-vector product
// (element forces, sum of dP).
constexpr int kFsElemPerBlock = kFsThreads / 8;
constexpr double kGaussSame = 0.5 + kGaussOff;   // weight of a node on the point's side of an axis
constexpr double kGaussOther = 0.5 - kGaussOff;  // on the other side

// T_j[d]: gradient of the shape function of label j along axis d at the point labelled 0
__host__ __device__ constexpr double tgrad(int j, int d) {
  const double fx = (j & 1) ? kGaussOther : kGaussSame, fy = (j & 2) ? kGaussOther : kGaussSame,
               fz = (j & 4) ? kGaussOther : kGaussSame;
  return d == 0 ? ((j & 1) ? fy * fz : -fy * fz) : (d == 1 ? ((j & 2) ? fx * fz : -fx * fz) : ((j & 4) ? fx * fy : -fx * fy));
}

// v[i ^ XM] inside each group of 32 lanes (XM < 32): ds_swizzle in its bit mode (and-mask 0x1f, or-mask 0,
// xor-mask XM), which needs no lane-index arithmetic, unlike a bpermute
template <int XM>
__device__ __forceinline__ double swz_xor(double v) {
  static_assert(XM > 0 && XM < 32, "swizzle partners lie in the same 32 lanes");
  constexpr int kPattern = 0x1F | (XM << 10);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), kPattern);
  const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), kPattern);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// sum over the 64 lanes of a wavefront, the same fixed xor butterfly in every lane
__device__ __forceinline__ double wave_sum(double v) {
  v += __shfl_xor(v, 32, 64);
  v += swz_xor<16>(v);
  v += swz_xor<8>(v);
  v += swz_xor<4>(v);
  v += swz_xor<2>(v);
  v += swz_xor<1>(v);
  return v;
}

// HASV: the product's direction field v is given (MODE >= 2); without it only dFbar enters dF. A template
// parameter rather than a test of a.v inside the node loop, where the test made every load of v wait for all
// loads in flight.
template <int MODE, bool HASV = false>
__global__ void __launch_bounds__(kFsThreads) fs_element_kernel(const FsArgs a) {
  const int g = threadIdx.x & 7;
  // 32-bit indexing: the host limits cells to 24 N < 2^31 (the force buffer's extent)
  const int n = (int)a.n;
  const int e = (int)blockIdx.x * kFsElemPerBlock + (int)(threadIdx.x >> 3);
  const bool live = e < n;
  double acc[kFsParts];
#pragma unroll
  for (int q = 0; q < kFsParts; ++q) acc[q] = 0.0;
  const double sig[3] = {(g & 1) ? -1.0 : 1.0, (g & 2) ? -1.0 : 1.0, (g & 4) ? -1.0 : 1.0};
  double P[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = 0.0;
  if (live) {
    const int row = (int)a.dz.div((unsigned)e);  // i ny + j
    const int k = e - row * a.nz;
    const int i = (int)a.dy.div((unsigned)row);
    const int j = row - i * a.ny;
    const int i1 = i + 1 == a.nx ? 0 : i + 1, j1 = j + 1 == a.ny ? 0 : j + 1, k1 = k + 1 == a.nz ? 0 : k + 1;
    // offsets (in values) of the element's four node rows (x, y corner) and of its two z positions
    const int r00 = row * a.nz * 3, r10 = (i1 * a.ny + j) * a.nz * 3;
    const int r01 = (i * a.ny + j1) * a.nz * 3, r11 = (i1 * a.ny + j1) * a.nz * 3;
    const int zoff[2] = {3 * k, 3 * k1};
    double F[3][3], dF[3][3];
#pragma unroll
    for (int ci = 0; ci < 3; ++ci)
#pragma unroll
      for (int dj = 0; dj < 3; ++dj) {
        F[ci][dj] = a.Fbar[ci * 3 + dj];
        dF[ci][dj] = MODE >= 2 ? a.dFbar[ci * 3 + dj] : 0.0;
      }
    // this point's weight along each axis of a node on the low (index 0) or high (1) side of the element:
    // kGaussSame when node and point sit on the same side
    const double wx[2] = {(g & 1) ? kGaussOther : kGaussSame, (g & 1) ? kGaussSame : kGaussOther};
    const double wy[2] = {(g & 2) ? kGaussOther : kGaussSame, (g & 2) ? kGaussSame : kGaussOther};
    const double wz[2] = {(g & 4) ? kGaussOther : kGaussSame, (g & 4) ? kGaussSame : kGaussOther};
    // the loads go node by node, the same node for the 8 lanes of an element (one address per element and
    // instruction). Node c = x + 2 y + 4 z (bits of its corner) has at this point the gradient
    // {(x ? 1 : -1) wy[y] wz[z], (y ? 1 : -1) wx[x] wz[z], (z ? 1 : -1) wx[x] wy[y]}; the loop runs over the
    // four (y, z) pairs of nodes x = 0, 1, whose y and z are uniform over the wavefront (scalar selects), with
    // two nodes' loads in flight (unrolling all eight holds 48 loaded values in registers).
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
      const bool hy = t & 1, hz = t >> 1;
      const double fy = hy ? wy[1] : wy[0], fz = hz ? wz[1] : wz[0];
      const double dyz = fy * fz;
      const double fzs = hy ? fz : -fz, fys = hz ? fy : -fy;
      const int zo = hz ? zoff[1] : zoff[0];
      const int nd[2] = {(hy ? r01 : r00) + zo, (hy ? r11 : r10) + zo};
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const double d[3] = {x ? dyz : -dyz, wx[x] * fzs, wx[x] * fys};
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
          const double wv = a.w[nd[x] + ci];
#pragma unroll
          for (int dj = 0; dj < 3; ++dj) F[ci][dj] += wv * d[dj];
          if constexpr (MODE >= 2 && HASV) {
            const double vv = a.v[nd[x] + ci];
#pragma unroll
            for (int dj = 0; dj < 3; ++dj) dF[ci][dj] += vv * d[dj];
          }
        }
      }
    }
    const double sc = a.sg[e];
    const double gam = a.sg[n + e];
    if (MODE == 3) {
      // d/drho_e of wq s_e sum_g P(F_g; gamma_e) : dF_g, this lane's point: summed over the 8 lanes below
      double Pv[3][3], Pgam[3][3];
      eff_P_dgamma(F, a.m, gam, Pv, Pgam);
      double t = 0.0, u = 0.0;
#pragma unroll
      for (int ci = 0; ci < 3; ++ci)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          t += Pv[ci][dj] * dF[ci][dj];
          u += Pgam[ci][dj] * dF[ci][dj];
        }
      acc[0] = a.sg[2 * n + e] * t + sc * a.sg[3 * n + e] * u;
    } else if (MODE == 2) {
      eff_dP(F, a.m, gam, dF, P);
    } else {
      acc[0] = sc * eff_psi_P(F, a.m, gam, P);
    }
    if (MODE == 1 || MODE == 2) {
#pragma unroll
      for (int ci = 0; ci < 3; ++ci)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          P[ci][dj] *= sc;
          acc[1 + ci * 3 + dj] = P[ci][dj];
        }
    }
  }
  if (MODE == 3) {
    // per-element sum over the 8 Gauss points (fixed butterfly); no block partials
    double t = acc[0];
    t += swz_xor<4>(t);
    t += swz_xor<2>(t);
    t += swz_xor<1>(t);
    if (live && g == 0) a.fe[e] = t;
    return;
  }
  if (MODE != 0) {
    // P (.) sigma: the forces in labelled form use the constants T_j (zero for lanes past the end)
    double Ps[3][3];
#pragma unroll
    for (int ci = 0; ci < 3; ++ci)
#pragma unroll
      for (int dj = 0; dj < 3; ++dj) Ps[ci][dj] = P[ci][dj] * sig[dj];
    auto contrib = [&](int jl, int ci) {
      return Ps[ci][0] * tgrad(jl, 0) + Ps[ci][1] * tgrad(jl, 1) + Ps[ci][2] * tgrad(jl, 2);
    };
    double h4[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) h4[q][ci] = contrib(q, ci) + swz_xor<4>(contrib(q + 4, ci));
    double h2[2][3];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) h2[q][ci] = h4[q][ci] + swz_xor<2>(h4[q + 2][ci]);
    double h1[3];
#pragma unroll
    for (int ci = 0; ci < 3; ++ci) h1[ci] = h2[0][ci] + swz_xor<1>(h2[1][ci]);
    if (live) {
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) a.fe[(g * 3 + ci) * n + e] = h1[ci];
    }
  }
  // partial sums of the block: a fixed xor butterfly per wavefront for every value, then one barrier and a
  // fixed-order sum over the wavefronts; stored value-major (part[q][block]) for coalesced final sums
  // only the sums the caller asked for (a.pmask; the others are stored as zero)
  constexpr int nq = MODE == 0 ? 1 : kFsParts;
  __shared__ double redq[kFsThreads / 64][kFsParts];
#pragma unroll
  for (int q = 0; q < nq; ++q) {
    if (!((a.pmask >> q) & 1u)) continue;  // uniform over the launch
    const double v = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) redq[threadIdx.x >> 6][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < nq) {
    double t = 0.0;
    if ((a.pmask >> threadIdx.x) & 1u) {
#pragma unroll
      for (int w = 0; w < kFsThreads / 64; ++w) t += redq[w][threadIdx.x];
    }
    a.part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// sg = (wq s_e, gamma_e, wq ds_e/drho_e, dgamma_e/drho_e) per element: s = emin + (1 - emin) clip01(rho)^p
// (mech.py:104-105) and the void interpolation factor gamma (mech.py:107-112; 1 without it), with the
// derivative convention of the reference's clip01 (1 on [0, 1] including the bounds, 0 outside; mech.py:31-35)
__global__ void __launch_bounds__(kFsThreads) fs_prepare_kernel(const double* __restrict__ rho, FsMat m, double wq,
                                                                int64_t n, double* __restrict__ sg) {
  const int64_t e = (int64_t)blockIdx.x * kFsThreads + threadIdx.x;
  if (e >= n) return;
  const double r = rho[e];
  const double c = clip01(r);
  const double dc = (r > 1.0 || r < 0.0) ? 0.0 : 1.0;
  const double vp = pow(c, m.p);
  const double dvp = dc * m.p * pow(c, m.p - 1.0);
  double gam = 1.0, dgam = 0.0;
  if (m.wls) {
    const double b = m.wls_b, r0 = m.wls_r0;
    const double den = tanh(b * r0) + tanh(b * (1.0 - r0));
    const double th = tanh(b * (vp - r0));
    gam = (tanh(b * r0) + th) / den;
    dgam = b * (1.0 - th * th) / den * dvp;
  }
  sg[e] = wq * (m.emin + (1.0 - m.emin) * vp);
  sg[n + e] = gam;
  sg[2 * n + e] = wq * (1.0 - m.emin) * dvp;
  sg[3 * n + e] = dgam;
}

// out(node, c) = sum over the node's 8 elements of their force on it (fixed order: corner 0..7)
__global__ void __launch_bounds__(kFsThreads) fs_gather_kernel(const FsArgs a, double* __restrict__ out) {
  const int n = (int)a.n;  // 24 n < 2^31 (host-checked)
  const int nd = (int)blockIdx.x * kFsThreads + (int)threadIdx.x;
  if (nd >= n) return;
  const int row = (int)a.dz.div((unsigned)nd);
  const int k = nd - row * a.nz;
  const int i = (int)a.dy.div((unsigned)row);
  const int j = row - i * a.ny;
  const int i0 = i == 0 ? a.nx - 1 : i - 1, j0 = j == 0 ? a.ny - 1 : j - 1, k0 = k == 0 ? a.nz - 1 : k - 1;
  double o[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    // the element whose corner c is this node starts one voxel back along each set bit of c
    const int ii = (c & 1) ? i0 : i, jj = (c & 2) ? j0 : j, kk = (c & 4) ? k0 : k;
    const int el = (ii * a.ny + jj) * a.nz + kk;
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] += a.fe[(c * 3 + d) * n + el];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) out[nd * 3 + d] = o[d];
}

// Mean removal of an (N, 3) field in two launches (mech.py:236-239, the w part of Mixed.proj): per-block
// component sums over a grid-stride sweep of at most kProjBlocks blocks (one node per thread and step), then a
// subtraction pass whose every block first forms the three means from those partials, each component by one
// wavefront in a fixed order (lane-strided sums, then a fixed butterfly), so all blocks subtract bitwise the
// same values.
constexpr int kProjBlocks = 512;

__global__ void __launch_bounds__(kFsThreads) fs_sum3_kernel(const double* __restrict__ w, int n,
                                                             double* __restrict__ part) {
  __shared__ double red[kFsThreads / 64];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int nd = (int)blockIdx.x * kFsThreads + (int)threadIdx.x; nd < n; nd += (int)gridDim.x * kFsThreads) {
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += w[nd * 3 + c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double t = block_sum(acc[c], red);
    if (threadIdx.x == 0) part[c * gridDim.x + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kFsThreads) fs_sub3_kernel(double* __restrict__ w, int n,
                                                             const double* __restrict__ part, int nb, double inv_n) {
  __shared__ double mean[3];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wv < 3) {
    double acc = 0.0;
#pragma unroll 8
    for (int b = lane; b < nb; b += 64) acc += part[wv * nb + b];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) mean[wv] = acc * inv_n;
  }
  __syncthreads();
  const int len = 3 * n;
  for (int i = (int)blockIdx.x * kFsThreads + (int)threadIdx.x; i < len; i += (int)gridDim.x * kFsThreads)
    w[i] -= mean[i % 3];
}

// The macro unknowns of a mixed solve: k free components of Fbar at row-major positions idx (host values,
// passed by value). scatter: dst (9) = the 3x3 matrix with src[j] at idx[j] and zeros elsewhere; gather:
// dst[j] = src[idx[j]].
struct FsIdx {
  int k;
  int i[9];
};

__global__ void fs_macro_scatter_kernel(const FsIdx ix, const double* __restrict__ src, double* __restrict__ dst) {
  const int t = threadIdx.x;
  if (t < 9) {
    double v = 0.0;
    for (int j = 0; j < ix.k; ++j)
      if (ix.i[j] == t) v = src[j];
    dst[t] = v;
  }
}

__global__ void fs_macro_gather_kernel(const FsIdx ix, const double* __restrict__ src, double* __restrict__ dst) {
  const int t = threadIdx.x;
  if (t < ix.k) dst[t] = src[ix.i[t]];
}

// Dot products of the CG without a reduction launch of their own: the subtraction pass of a projection writes
// kDotParts per-block partials of y . w (w after the subtraction, its (N, 3) part), and the kernel that needs the
// value sums them itself, every block in the same fixed order (two values per lane, then a fixed butterfly), and
// adds the k macro terms in order.
constexpr int kDotParts = 256;

__device__ __forceinline__ double dot_total(const double* __restrict__ part, const double* __restrict__ a,
                                            const double* __restrict__ b, int k) {
  const int lane = threadIdx.x & 63;
  static_assert(kDotParts == 256, "four partials per lane");
  double acc = (part[lane] + part[lane + 64]) + (part[lane + 128] + part[lane + 192]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  for (int j = 0; j < k; ++j) acc += a[j] * b[j];
  return acc;
}

__global__ void __launch_bounds__(kFsThreads) fs_sub3_dot_kernel(double* __restrict__ w, int n,
                                                                 const double* __restrict__ part, int nb, double inv_n,
                                                                 const double* __restrict__ y,
                                                                 double* __restrict__ dpart) {
  __shared__ double mean[3];
  __shared__ double red[kFsThreads / 64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wv < 3) {
    double acc = 0.0;
#pragma unroll 8
    for (int b = lane; b < nb; b += 64) acc += part[wv * nb + b];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) mean[wv] = acc * inv_n;
  }
  __syncthreads();
  const int len = 3 * n;
  double acc = 0.0;
  for (int i = (int)blockIdx.x * kFsThreads + (int)threadIdx.x; i < len; i += (int)gridDim.x * kFsThreads) {
    const double v = w[i] - mean[i % 3];
    w[i] = v;
    acc += y[i] * v;
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) dpart[blockIdx.x] = t;
}

// x_out = x + alpha p, r_out = r - alpha Ap with alpha = rz / pAp (0 if pAp <= 0), pAp from its partials and the
// macro terms (the last k entries); block 0 stores pAp; per-block partials of |r_out|^2
__global__ void __launch_bounds__(kFsThreads) fs_cg_update_parts_kernel(int len, int k, double* __restrict__ x_out,
                                                                        double* __restrict__ r_out,
                                                                        const double* __restrict__ x,
                                                                        const double* __restrict__ r,
                                                                        const double* __restrict__ p,
                                                                        const double* __restrict__ Ap,
                                                                        const double* __restrict__ rz,
                                                                        const double* __restrict__ pap_part,
                                                                        double* __restrict__ pap_out,
                                                                        double* __restrict__ rn2_part) {
  __shared__ double red[kFsThreads / 64];
  __shared__ double al_sh;
  if (threadIdx.x < 64) {
    const double pa = dot_total(pap_part, p + (len - k), Ap + (len - k), k);
    if (threadIdx.x == 0) {
      al_sh = pa <= 0.0 ? 0.0 : *rz / pa;
      if (blockIdx.x == 0) *pap_out = pa;
    }
  }
  __syncthreads();
  const double al = al_sh;
  const int i = (int)blockIdx.x * kFsThreads + (int)threadIdx.x;
  double v = 0.0;
  if (i < len) {
    x_out[i] = x[i] + al * p[i];
    const double rn = r[i] - al * Ap[i];
    r_out[i] = rn;
    v = rn * rn;
  }
  const double t = block_sum(v, red);
  if (threadIdx.x == 0) rn2_part[blockIdx.x] = t;
}

// p_out = z + beta p with beta = rz_new / rz_old (rz_old = 0 taken as 1), rz_new = r . z from its partials and the
// macro terms; block 0 stores rz_new and |r|^2 summed from the update's per-block partials by all its threads
// (fixed order: a thread's partials in turn, then the block sum)
__global__ void __launch_bounds__(kFsThreads) fs_cg_direction_parts_kernel(int len, int k, double* __restrict__ p_out,
                                                                           const double* __restrict__ z,
                                                                           const double* __restrict__ p,
                                                                           const double* __restrict__ r,
                                                                           const double* __restrict__ rz_old,
                                                                           const double* __restrict__ rz_part,
                                                                           double* __restrict__ rz_new,
                                                                           const double* __restrict__ rn2_part,
                                                                           int nrn2, double* __restrict__ rn2_out) {
  __shared__ double beta_sh;
  __shared__ double red[kFsThreads / 64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wv == 0) {
    const double rz1 = dot_total(rz_part, r + (len - k), z + (len - k), k);
    if (lane == 0) {
      const double ro = *rz_old;
      beta_sh = rz1 / (ro == 0.0 ? 1.0 : ro);
      if (blockIdx.x == 0) *rz_new = rz1;
    }
  }
  if (blockIdx.x == 0) {  // uniform over the block
    double acc = 0.0;
#pragma unroll 8
    for (int b = threadIdx.x; b < nrn2; b += kFsThreads) acc += rn2_part[b];
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) *rn2_out = t;
  }
  __syncthreads();
  const double beta = beta_sh;
  const int i = (int)blockIdx.x * kFsThreads + (int)threadIdx.x;
  if (i < len) p_out[i] = z[i] + beta * p[i];
}

// Third-order elastic constants of the cell from the linear correctors chi_I (6 fields (N, 3)), per
// Gauss point H_I = e_I + grad chi_I with e_I the symmetric Voigt unit strain (engineering shear: 1/2
// off the diagonal), at F = I (toec, mech.py:454-490), with f(s) = <psi_eff(I + sum_I s_I H_I)>:
//   d2f_IJ  = sum_e wq s_e sum_g D2psi_eff(I)[H_I, H_J],  D2psi_eff(I) = D2psi_lin (the neo-Hookean tangent at I
//            is the linear one, so the void interpolation drops out): lam tr a tr b + mu (a : b + a : b^T);
//   d3f_IJK = sum_e wq s_e gamma_e^3 sum_g D3psi_nl(I)[H_I, H_J, H_K] (psi_lin is quadratic), with
//            D3psi_neo(I)[a, b, c] = -mu (tr abc + tr acb) - lam (tr a tr bc + tr b tr ac + tr c tr ab), 0 linear.
// Eight lanes per element (one per Gauss point); each of the 21 + 56 unique values is block-summed in a
// fixed order into part[block][77] (pairs I <= J first, then triples I <= J <= K).
constexpr int kToecVals = 77;

__device__ __forceinline__ double tr_prod3(const double (&a)[3][3], const double (&b)[3][3], const double (&c)[3][3]) {
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double ab = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
      t += ab * c[j][i];
    }
  return t;
}

__global__ void __launch_bounds__(kFsThreads) fs_toec_kernel(const FsArgs a, const double* __restrict__ chi,
                                                             double* __restrict__ part) {
  __shared__ double red[kFsThreads / 64];
  const int g = threadIdx.x & 7;
  const int n = (int)a.n;
  const int e = (int)blockIdx.x * kFsElemPerBlock + (int)(threadIdx.x >> 3);
  const bool live = e < n;
  double H[6][3][3];
#pragma unroll
  for (int I = 0; I < 6; ++I)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) H[I][i][j] = 0.0;
  double ws = 0.0, wg3 = 0.0;
  if (live) {
    const int row = (int)a.dz.div((unsigned)e);
    const int k = e - row * a.nz;
    const int i = (int)a.dy.div((unsigned)row);
    const int j = row - i * a.ny;
    const int i1 = i + 1 == a.nx ? 0 : i + 1, j1 = j + 1 == a.ny ? 0 : j + 1, k1 = k + 1 == a.nz ? 0 : k + 1;
    for (int c = 0; c < 8; ++c) {
      const int ii = (c & 1) ? i1 : i, jj = (c & 2) ? j1 : j, kk = (c & 4) ? k1 : k;
      const int nd = (ii * a.ny + jj) * a.nz + kk;
      const int jl = c ^ g;
      const double fx = (jl & 1) ? kGaussOther : kGaussSame, fy = (jl & 2) ? kGaussOther : kGaussSame,
                   fz = (jl & 4) ? kGaussOther : kGaussSame;
      const double d[3] = {((c & 1) ? fy : -fy) * fz, ((c & 2) ? fx : -fx) * fz, ((c & 4) ? fx : -fx) * fy};
#pragma unroll
      for (int I = 0; I < 6; ++I)
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
          const double x = chi[((size_t)I * n + nd) * 3 + ci];
#pragma unroll
          for (int dj = 0; dj < 3; ++dj) H[I][ci][dj] += x * d[dj];
        }
    }
    // + e_I: Voigt order xx yy zz yz xz xy
    H[0][0][0] += 1.0;
    H[1][1][1] += 1.0;
    H[2][2][2] += 1.0;
    H[3][1][2] += 0.5;
    H[3][2][1] += 0.5;
    H[4][0][2] += 0.5;
    H[4][2][0] += 0.5;
    H[5][0][1] += 0.5;
    H[5][1][0] += 0.5;
    ws = a.sg[e];
    const double gam = a.sg[n + e];
    wg3 = a.m.neo ? ws * gam * gam * gam : 0.0;
  }
  double tr[6];
#pragma unroll
  for (int I = 0; I < 6; ++I) tr[I] = H[I][0][0] + H[I][1][1] + H[I][2][2];
  const double lam = a.m.lam, mu = a.m.mu;
  int q = 0;
  for (int I = 0; I < 6; ++I)
    for (int J = I; J < 6; ++J, ++q) {
      double dd = 0.0;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dd += H[I][r][c] * (H[J][r][c] + H[J][c][r]);
      const double t = block_sum(ws * (lam * tr[I] * tr[J] + mu * dd), red);
      if (threadIdx.x == 0) part[(size_t)blockIdx.x * kToecVals + q] = t;
      __syncthreads();
    }
  for (int I = 0; I < 6; ++I)
    for (int J = I; J < 6; ++J)
      for (int K = J; K < 6; ++K, ++q) {
        double v = 0.0;
        if (wg3 != 0.0) {
          double tIJ = 0.0, tIK = 0.0, tJK = 0.0;
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              tIJ += H[I][r][c] * H[J][c][r];
              tIK += H[I][r][c] * H[K][c][r];
              tJK += H[J][r][c] * H[K][c][r];
            }
          const double d3 = -mu * (tr_prod3(H[I], H[J], H[K]) + tr_prod3(H[I], H[K], H[J])) -
                            lam * (tr[I] * tJK + tr[J] * tIK + tr[K] * tIJ);
          v = wg3 * d3;
        }
        const double t = block_sum(v, red);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * kToecVals + q] = t;
        __syncthreads();
      }
}

__global__ void __launch_bounds__(kFsThreads) fs_toec_reduce_kernel(const double* __restrict__ part, int nblk,
                                                                    double* __restrict__ out) {
  __shared__ double red[kFsThreads / 64];
  const int q = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kFsThreads) acc += part[(size_t)b * kToecVals + q];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) out[q] = t;
}

// CG vector updates of the Newton-CG driver (Mixed._cg, mech.py:266-283) on the flat unknowns z = (w, Fbar_free):
// alpha = rz / pAp (0 when pAp <= 0: the negative-curvature exit keeps x), x' = x + alpha p, r' = r - alpha Ap,
// with the block partial sums of r'^2; and p' = z + (rz_new / rz_old) p (rz_old = 0 -> 1).
__global__ void __launch_bounds__(kFsThreads) fs_cg_update_kernel(int len, double* __restrict__ x_out,
                                                                  double* __restrict__ r_out, const double* __restrict__ x,
                                                                  const double* __restrict__ r,
                                                                  const double* __restrict__ p,
                                                                  const double* __restrict__ Ap,
                                                                  const double* __restrict__ rz,
                                                                  const double* __restrict__ pAp,
                                                                  double* __restrict__ part) {
  __shared__ double red[kFsThreads / 64];
  const double pa = *pAp;
  const double al = pa <= 0.0 ? 0.0 : *rz / pa;
  const int i = (int)blockIdx.x * kFsThreads + (int)threadIdx.x;
  double v = 0.0;
  if (i < len) {
    x_out[i] = x[i] + al * p[i];
    const double rn = r[i] - al * Ap[i];
    r_out[i] = rn;
    v = rn * rn;
  }
  const double t = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kFsThreads) fs_cg_direction_kernel(int len, double* __restrict__ p_out,
                                                                     const double* __restrict__ z,
                                                                     const double* __restrict__ p,
                                                                     const double* __restrict__ rz_new,
                                                                     const double* __restrict__ rz_old) {
  const double ro = *rz_old;
  const double beta = *rz_new / (ro == 0.0 ? 1.0 : ro);
  const int i = (int)blockIdx.x * kFsThreads + (int)threadIdx.x;
  if (i < len) p_out[i] = z[i] + beta * p[i];
}

// out[q] = sum over blocks of part[b][q], lane-strided then a fixed butterfly (deterministic)
constexpr int kFsReduceThreads = 1024;

__global__ void __launch_bounds__(kFsReduceThreads) fs_reduce_kernel(const double* __restrict__ part, int nblk,
                                                                     double* __restrict__ out) {
  __shared__ double red[kFsReduceThreads / 64];
  const int q = blockIdx.x;
  const double* pq = part + (int64_t)q * nblk;  // value-major: the partials of value q are contiguous
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kFsReduceThreads) acc += pq[b];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) out[q] = t;
}

}  // namespace
}  // namespace nlfh

using namespace nlfh;

extern "C" {

long long nlfh_fs_work_doubles(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t nblk = (n + kFsElemPerBlock - 1) / kFsElemPerBlock;
  return (long long)(24 * n + nblk * kFsParts);
}

static int fs_setup(const int* shape, const double* mat, int flags, FsArgs& a) {
  if (shape == nullptr || mat == nullptr) {
    set_error("nlfh_fs: shape and mat are required");
    return NLFH_EINVAL;
  }
  const int nx = shape[0], ny = shape[1], nz = shape[2];
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("nlfh_fs: cell dimensions must be positive");
    return NLFH_EINVAL;
  }
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.n = (int64_t)nx * ny * nz;
  a.dz = FastDiv((unsigned)nz);
  a.dy = FastDiv((unsigned)ny);
  if (24 * a.n >= (int64_t)INT32_MAX) {  // the kernels index the (24, N) force buffer in 32 bits
    set_error("nlfh_fs: cell too large (24 N must stay below 2^31)");
    return NLFH_EINVAL;
  }
  a.m.lam = mat[0];
  a.m.mu = mat[1];
  a.m.emin = mat[2];
  a.m.p = mat[3];
  a.m.wls_b = mat[4];
  a.m.wls_r0 = mat[5];
  a.m.neo = (flags & 1) ? 1 : 0;
  a.m.wls = (flags & 2) ? 1 : 0;
  a.wq = 1.0 / (8.0 * (double)a.n);
  return NLFH_OK;
}

int nlfh_fs_prepare(const int* shape, const double* rho, const double* mat, int flags, double* sg, void* stream) {
  FsArgs a{};
  const int rc = fs_setup(shape, mat, flags, a);
  if (rc != NLFH_OK) return rc;
  if (rho == nullptr || sg == nullptr) {
    set_error("nlfh_fs_prepare: rho and sg are required");
    return NLFH_EINVAL;
  }
  const int64_t gblk = (a.n + kFsThreads - 1) / kFsThreads;
  hipLaunchKernelGGL(fs_prepare_kernel, dim3((unsigned)gblk), dim3(kFsThreads), 0, (hipStream_t)stream, rho, a.m, a.wq,
                     a.n, sg);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

}  // extern "C"

namespace nlfh {
namespace {

// The workspace of the no-allocation form of nlfh_fs_eval: the (24, N) element forces and the block partials
// of a cell of up to kNoAllocMaxVoxels voxels (32^3), in the library's own device memory. Host code reaches it
// only through the per-device address from hipGetSymbolAddress (the array's name is a host-side stand-in).
constexpr int64_t kNoAllocMaxVoxels = 32768;
constexpr int64_t kNoAllocDoubles =
    24 * kNoAllocMaxVoxels + ((kNoAllocMaxVoxels + kFsElemPerBlock - 1) / kFsElemPerBlock) * kFsParts;
__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the scratch array's address on the current device, looked up once per device
int scratch_on_current_device(double** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("nlfh_fs_eval_noalloc: device index out of range");
    return NLFH_EUNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      set_error(std::string("nlfh_fs_eval_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return NLFH_EHIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return NLFH_OK;
}

int fs_eval_impl(const int* shape, const double* sg, const double* mat, int flags, const double* w,
                 const double* Fbar, const double* v, const double* dFbar, int mode, double* grad, double* scalars,
                 double* work, int64_t max_voxels, void* stream) {
  FsArgs a{};
  const int rc = fs_setup(shape, mat, flags, a);
  if (rc != NLFH_OK) return rc;
  if (mode < 0 || mode > 3) {
    set_error("nlfh_fs_eval: mode must be 0 (energy), 1 (energy and gradient), 2 (Hessian-vector product) or 3 "
              "(density sensitivity)");
    return NLFH_EINVAL;
  }
  if (sg == nullptr || w == nullptr || Fbar == nullptr || work == nullptr || (mode <= 2 && scalars == nullptr) ||
      (mode >= 1 && grad == nullptr) || (mode >= 2 && dFbar == nullptr)) {
    set_error("nlfh_fs_eval: a required pointer is null");
    return NLFH_EINVAL;
  }
  if (max_voxels > 0 && a.n > max_voxels) {
    set_error("nlfh_fs_eval_noalloc: the cell has " + std::to_string(a.n) + " voxels, more than the " +
              std::to_string(max_voxels) + " its static workspace holds (use nlfh_fs_eval with a workspace)");
    return NLFH_EUNSUPPORTED;
  }
  a.sg = sg;
  a.w = w;
  a.v = mode >= 2 ? v : nullptr;
  // flags bit 2: form only the scalars named in bits 8..17 (bit 8 + q: scalar q)
  a.pmask = (flags & 4) ? ((unsigned)flags >> 8) & 0x3FFu : 0x3FFu;
  for (int q = 0; q < 9; ++q) a.Fbar[q] = Fbar[q];
  a.dFbar = mode >= 2 ? dFbar : nullptr;
  const int64_t nblk = (a.n + kFsElemPerBlock - 1) / kFsElemPerBlock;  // element blocks (partials)
  const int64_t gblk = (a.n + kFsThreads - 1) / kFsThreads;             // gather blocks
  a.fe = work;
  a.part = work + 24 * a.n;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)nblk);
  if (mode == 3) {  // per-element output straight into grad (N values)
    a.fe = grad;
    if (a.v != nullptr) hipLaunchKernelGGL((fs_element_kernel<3, true>), grid, dim3(kFsThreads), 0, st, a);
    else hipLaunchKernelGGL((fs_element_kernel<3, false>), grid, dim3(kFsThreads), 0, st, a);
    HIP_TRY(hipGetLastError());
    return NLFH_OK;
  }
  if (mode == 0) hipLaunchKernelGGL(fs_element_kernel<0>, grid, dim3(kFsThreads), 0, st, a);
  else if (mode == 1) hipLaunchKernelGGL(fs_element_kernel<1>, grid, dim3(kFsThreads), 0, st, a);
  else if (a.v != nullptr) hipLaunchKernelGGL((fs_element_kernel<2, true>), grid, dim3(kFsThreads), 0, st, a);
  else hipLaunchKernelGGL((fs_element_kernel<2, false>), grid, dim3(kFsThreads), 0, st, a);
  HIP_TRY(hipGetLastError());
  if (mode >= 1) {
    hipLaunchKernelGGL(fs_gather_kernel, dim3((unsigned)gblk), dim3(kFsThreads), 0, st, a, grad);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(fs_reduce_kernel, dim3(mode == 0 ? 1 : kFsParts), dim3(kFsReduceThreads), 0, st, a.part, (int)nblk,
                     scalars);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

}  // namespace
}  // namespace nlfh

extern "C" {

int nlfh_fs_eval(const int* shape, const double* sg, const double* mat, int flags, const double* w,
                 const double* Fbar, const double* v, const double* dFbar, int mode, double* grad, double* scalars,
                 double* work, void* stream) {
  return fs_eval_impl(shape, sg, mat, flags, w, Fbar, v, dFbar, mode, grad, scalars, work, 0, stream);
}

long long nlfh_fs_noalloc_max_voxels(void) { return (long long)kNoAllocMaxVoxels; }

int nlfh_fs_eval_noalloc(const int* shape, const double* sg, const double* mat, int flags, const double* w,
                         const double* Fbar, const double* v, const double* dFbar, int mode, double* grad,
                         double* scalars, void* stream) {
  double* work = nullptr;
  const int rc = scratch_on_current_device(&work);
  if (rc != NLFH_OK) return rc;
  return fs_eval_impl(shape, sg, mat, flags, w, Fbar, v, dFbar, mode, grad, scalars, work, kNoAllocMaxVoxels,
                      stream);
}

int nlfh_fs_toec(const int* shape, const double* sg, const double* mat, int flags, const double* chi, double* out,
                 double* work, void* stream) {
  FsArgs a{};
  const int rc = fs_setup(shape, mat, flags, a);
  if (rc != NLFH_OK) return rc;
  if (sg == nullptr || chi == nullptr || out == nullptr || work == nullptr) {
    set_error("nlfh_fs_toec: sg, chi, out and work are required");
    return NLFH_EINVAL;
  }
  if (6 * 3 * a.n >= (int64_t)INT32_MAX) {
    set_error("nlfh_fs_toec: cell too large");
    return NLFH_EINVAL;
  }
  a.sg = sg;
  const int nblk = (int)((a.n + kFsElemPerBlock - 1) / kFsElemPerBlock);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fs_toec_kernel, dim3(nblk), dim3(kFsThreads), 0, st, a, chi, work);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(fs_toec_reduce_kernel, dim3(kToecVals), dim3(kFsThreads), 0, st, work, nblk, out);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_fs_cg_update(int len, double* x_out, double* r_out, const double* x, const double* r, const double* p,
                      const double* Ap, const double* rz, const double* pAp, double* rn2, double* work, void* stream) {
  if (len < 1 || x_out == nullptr || r_out == nullptr || x == nullptr || r == nullptr || p == nullptr ||
      Ap == nullptr || rz == nullptr || pAp == nullptr || rn2 == nullptr || work == nullptr) {
    set_error("nlfh_fs_cg_update: a required pointer is null or len < 1");
    return NLFH_EINVAL;
  }
  const int nb = (len + kFsThreads - 1) / kFsThreads;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fs_cg_update_kernel, dim3(nb), dim3(kFsThreads), 0, st, len, x_out, r_out, x, r, p, Ap, rz, pAp,
                     work);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(fs_reduce_kernel, dim3(1), dim3(kFsReduceThreads), 0, st, work, nb, rn2);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_fs_cg_direction(int len, double* p_out, const double* z, const double* p, const double* rz_new,
                         const double* rz_old, void* stream) {
  if (len < 1 || p_out == nullptr || z == nullptr || p == nullptr || rz_new == nullptr || rz_old == nullptr) {
    set_error("nlfh_fs_cg_direction: a required pointer is null or len < 1");
    return NLFH_EINVAL;
  }
  const int nb = (len + kFsThreads - 1) / kFsThreads;
  hipLaunchKernelGGL(fs_cg_direction_kernel, dim3(nb), dim3(kFsThreads), 0, (hipStream_t)stream, len, p_out, z, p,
                     rz_new, rz_old);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_fs_project(const int* shape, double* w, double* work, void* stream) {
  FsArgs a{};
  const double mat[6] = {0, 0, 0, 1, 0, 0};
  const int rc = fs_setup(shape, mat, 0, a);
  if (rc != NLFH_OK) return rc;
  if (w == nullptr || work == nullptr) {
    set_error("nlfh_fs_project: w and work are required");
    return NLFH_EINVAL;
  }
  const int n = (int)a.n;
  const int nodes_blk = (n + kFsThreads - 1) / kFsThreads;
  const int nb = nodes_blk < kProjBlocks ? nodes_blk : kProjBlocks;  // partials: 3 nb <= 3 ceil(N / 256)
  const int vals_blk = (3 * n + kFsThreads - 1) / kFsThreads;
  const int ns = vals_blk < 4 * kProjBlocks ? vals_blk : 4 * kProjBlocks;
  double* part = work;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fs_sum3_kernel, dim3(nb), dim3(kFsThreads), 0, st, w, n, part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(fs_sub3_kernel, dim3(ns), dim3(kFsThreads), 0, st, w, n, part, nb, 1.0 / n);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

static int fs_macro_idx(int k, const int* idx, bool distinct, FsIdx& ix, const char* who) {
  if (k < 0 || k > 9 || (k > 0 && idx == nullptr)) {
    set_error(std::string(who) + ": k must be in [0, 9] and idx given for k > 0");
    return NLFH_EINVAL;
  }
  ix.k = k;
  for (int j = 0; j < 9; ++j) ix.i[j] = 0;
  for (int j = 0; j < k; ++j) {
    if (idx[j] < 0 || idx[j] > 8) {
      set_error(std::string(who) + ": idx values must be in [0, 8]");
      return NLFH_EINVAL;
    }
    if (distinct)
      for (int q = 0; q < j; ++q)
        if (idx[q] == idx[j]) {
          set_error(std::string(who) + ": idx values must be distinct");
          return NLFH_EINVAL;
        }
    ix.i[j] = idx[j];
  }
  return NLFH_OK;
}

int nlfh_fs_macro_scatter(int k, const int* idx, const double* src, double* dst9, void* stream) {
  FsIdx ix{};
  const int rc = fs_macro_idx(k, idx, true, ix, "nlfh_fs_macro_scatter");
  if (rc != NLFH_OK) return rc;
  if (dst9 == nullptr || (k > 0 && src == nullptr)) {
    set_error("nlfh_fs_macro_scatter: dst9 (and src for k > 0) are required");
    return NLFH_EINVAL;
  }
  hipLaunchKernelGGL(fs_macro_scatter_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, ix, src, dst9);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_fs_macro_gather(int k, const int* idx, const double* src9, double* dst, void* stream) {
  FsIdx ix{};
  const int rc = fs_macro_idx(k, idx, false, ix, "nlfh_fs_macro_gather");
  if (rc != NLFH_OK) return rc;
  if (k == 0) return NLFH_OK;
  if (src9 == nullptr || dst == nullptr) {
    set_error("nlfh_fs_macro_gather: src9 and dst are required");
    return NLFH_EINVAL;
  }
  hipLaunchKernelGGL(fs_macro_gather_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, ix, src9, dst);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

long long nlfh_fs_dot_part_doubles(void) { return (long long)(kDotParts + 3 * kProjBlocks); }

int nlfh_fs_project_dot(const int* shape, double* w, const double* y, double* part, void* stream) {
  FsArgs a{};
  const double mat[6] = {0, 0, 0, 1, 0, 0};
  const int rc = fs_setup(shape, mat, 0, a);
  if (rc != NLFH_OK) return rc;
  if (w == nullptr || y == nullptr || part == nullptr) {
    set_error("nlfh_fs_project_dot: w, y and part are required");
    return NLFH_EINVAL;
  }
  const int n = (int)a.n;
  const int nodes_blk = (n + kFsThreads - 1) / kFsThreads;
  const int nb = nodes_blk < kProjBlocks ? nodes_blk : kProjBlocks;
  double* sums = part + kDotParts;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(fs_sum3_kernel, dim3(nb), dim3(kFsThreads), 0, st, w, n, sums);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(fs_sub3_dot_kernel, dim3(kDotParts), dim3(kFsThreads), 0, st, w, n, sums, nb, 1.0 / n, y, part);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_fs_cg_update_parts(int len, int k, double* x_out, double* r_out, const double* x, const double* r,
                            const double* p, const double* Ap, const double* rz, const double* pap_part,
                            double* pap_out, double* rn2_part, void* stream) {
  if (len < 1 || k < 0 || k > len || x_out == nullptr || r_out == nullptr || x == nullptr || r == nullptr ||
      p == nullptr || Ap == nullptr || rz == nullptr || pap_part == nullptr || pap_out == nullptr ||
      rn2_part == nullptr) {
    set_error("nlfh_fs_cg_update_parts: a required pointer is null, len < 1 or k outside [0, len]");
    return NLFH_EINVAL;
  }
  const int nb = (len + kFsThreads - 1) / kFsThreads;
  hipLaunchKernelGGL(fs_cg_update_parts_kernel, dim3(nb), dim3(kFsThreads), 0, (hipStream_t)stream, len, k, x_out,
                     r_out, x, r, p, Ap, rz, pap_part, pap_out, rn2_part);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

int nlfh_fs_cg_direction_parts(int len, int k, double* p_out, const double* z, const double* p, const double* r,
                               const double* rz_old, const double* rz_part, double* rz_new, const double* rn2_part,
                               double* rn2_out, void* stream) {
  if (len < 1 || k < 0 || k > len || p_out == nullptr || z == nullptr || p == nullptr || r == nullptr ||
      rz_old == nullptr || rz_part == nullptr || rz_new == nullptr || rn2_part == nullptr || rn2_out == nullptr ||
      rz_new == rz_old) {
    set_error("nlfh_fs_cg_direction_parts: a required pointer is null, rz_new is rz_old, len < 1 or k outside "
              "[0, len]");
    return NLFH_EINVAL;
  }
  const int nb = (len + kFsThreads - 1) / kFsThreads;
  hipLaunchKernelGGL(fs_cg_direction_parts_kernel, dim3(nb), dim3(kFsThreads), 0, (hipStream_t)stream, len, k, p_out,
                     z, p, r, rz_old, rz_part, rz_new, rn2_part, nb, rn2_out);
  HIP_TRY(hipGetLastError());
  return NLFH_OK;
}

}  // extern "C"
