// C-ABI of the library (include/nlfh.h): plans, launch tables, the preconditioner symbol, and the
// PCG / iterative-refinement drivers. Host code; every device step is a kernel of this library.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "nlfh.h"
#include "nlfh_kernels.h"

namespace nlfh {
thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
}  // namespace nlfh

using namespace nlfh;

namespace {
// Progress lines on stderr from the solver loops, at most one per 20 s, when NLFH_PROGRESS is set
// (long solves, e.g. many 128^3 cells at high contrast, would otherwise be silent for minutes).
bool progress_on() {
  static const bool on = [] {
    const char* e = std::getenv("NLFH_PROGRESS");
    return e != nullptr && e[0] != '\0' && e[0] != '0';
  }();
  return on;
}
template <typename... A>
void progress(const char* fmt, A... args) {
  if (!progress_on()) return;
  // at most one line per 20 s over all threads (plans may solve concurrently on their own threads)
  using clk = std::chrono::steady_clock;
  static std::atomic<int64_t> last{clk::now().time_since_epoch().count()};
  const int64_t now = clk::now().time_since_epoch().count();
  int64_t prev = last.load();
  const int64_t gap = std::chrono::duration_cast<clk::duration>(std::chrono::seconds(20)).count();
  if (now - prev < gap || !last.compare_exchange_strong(prev, now)) return;
  std::fprintf(stderr, "[nlfh] ");
  std::fprintf(stderr, fmt, args...);
  std::fprintf(stderr, "\n");
  std::fflush(stderr);
}
}  // namespace

#define CHECK_ARG(cond, msg)        \
  do {                              \
    if (!(cond)) {                  \
      nlfh::set_error(msg);         \
      return NLFH_EINVAL;           \
    }                               \
  } while (0)

#define TRY(expr)                 \
  do {                            \
    int _rc = (expr);             \
    if (_rc != NLFH_OK) return _rc; \
  } while (0)

namespace {

constexpr int kChunk = 4096;     // voxels per reduction chunk
constexpr int kMaxRhs = 6;
constexpr size_t kLdsMax = 160 * 1024 - 1024;

struct Shape {
  int nx, ny, nz;
  int ty, xc, zpt;
  int64_t off_sym;  // in entries of 6 reals
  int len[3];       // FftLen index per axis
};

struct MvGroup {
  int shape;
  int begin, count;
  int max_ty;
};

// Preconditioner launch group: the cells of one shape (fast register-FFT path for cubic cells of
// side 16, 32, 48, 64 or 128, generic path otherwise).
struct PrecGroup {
  int shape;
  int fast;  // side N of the fast path, 0 for generic
  int plane_begin, plane_count;
  int line_begin[2], line_count[2];  // [0] float64 line items, [1] float32
  int vec;  // every cell's field and spectrum offsets allow 16-byte vector access (staged passes)
  int lpb32 = 0;  // FP32 pass B positions per block on the fast path (0: the default)
  int ppa32 = 1;  // x-planes per block of the FP32 spectral pass A (1 or 2)
};

template <typename T>
struct Work {
  T* X = nullptr;
  T* R = nullptr;
  T* B = nullptr;
  T* P[2] = {nullptr, nullptr};
  T* Ap = nullptr;
  T* Z = nullptr;
  T* s = nullptr;
  cplx<T>* spec = nullptr;
  cplx<T>* rspec = nullptr;  // FP32 direction iteration: the residual's (z, y) spectrum
  double* st = nullptr;  // rz, alpha, beta, bn, rn (5 * ncells * M)
  int* ist = nullptr;    // it, done (2 * ncells)
  CgState cs{};
};

struct GraphKey {
  int dtype, m, iters;
  bool operator<(const GraphKey& o) const {
    if (dtype != o.dtype) return dtype < o.dtype;
    if (m != o.m) return m < o.m;
    return iters < o.iters;
  }
};

// digit-reversal map of the DIT transform with radices rad[0..nst-1]:
// position p holds x[idx(p)]
std::vector<int> digit_reversal(int n, const std::vector<int>& rad) {
  std::vector<int> idx(n);
  for (int p = 0; p < n; ++p) {
    int pp = p, len = n, mult = 1, v = 0;
    for (int s = (int)rad.size() - 1; s >= 0; --s) {
      len /= rad[s];
      const int q = pp / len;
      pp -= q * len;
      v += q * mult;
      mult *= rad[s];
    }
    idx[p] = v;
  }
  return idx;
}

bool factor(int n, std::vector<int>& rad) {
  rad.clear();
  int v = n;
  while (v % 4 == 0) { rad.push_back(4); v /= 4; }
  while (v % 2 == 0) { rad.push_back(2); v /= 2; }
  while (v % 3 == 0) { rad.push_back(3); v /= 3; }
  while (v % 5 == 0) { rad.push_back(5); v /= 5; }
  return v == 1 && (int)rad.size() <= NLFH_FFT_MAX_STAGES;
}

}  // namespace

struct nlfh_plan {
  int device = 0;
  int ncells = 0, m_layout = 6, M = 6;
  double E = 1.0, nu = 0.3, lam = 0, mu = 0;
  std::vector<CellDesc> cells;
  std::vector<int> cell_shape;
  std::vector<Shape> shapes;
  int64_t tot_vox = 0, tot_spec = 0, tot_sym = 0;

  std::vector<FftLen> lens;
  std::vector<MvItem> mv;
  std::vector<MvGroup> groups;
  std::vector<PlaneItem> planes;
  std::vector<LineItem> lines64, lines32;
  std::vector<PrecGroup> pgroups;
  double* d_symsoa64 = nullptr;  // fast shapes: component-planar symbol [6][nx][ny][nzh] / N
  float* d_symsoa32 = nullptr;
  std::vector<int64_t> symsoa_off;  // per shape (fast shapes only)
  std::vector<ChunkItem> chunks;
  int rpb64 = 1, rpb32 = 1;
  size_t lds_plane64 = 0, lds_plane32 = 0, lds_line64 = 0, lds_line32 = 0;

  // device tables
  CellDesc* d_cells = nullptr;
  FftLen* d_lens = nullptr;
  cplx<double>* d_tw64 = nullptr;
  cplx<float>* d_tw32 = nullptr;
  int* d_perm = nullptr;
  int* d_iperm = nullptr;
  int* d_shape_len = nullptr;
  double* d_sym64 = nullptr;
  float* d_sym32 = nullptr;
  std::vector<double> h_sym;  // unscaled real symbol (6 per frequency), for nlfh_symbol
  MvItem* d_mv = nullptr;
  PlaneItem* d_planes = nullptr;
  LineItem* d_lines64 = nullptr;
  LineItem* d_lines32 = nullptr;
  ChunkItem* d_chunks = nullptr;
  int* d_mv_range = nullptr;
  int* d_plane_range = nullptr;
  int* d_line_range32 = nullptr;  // per cell [begin, end) of its float32 line items
  double* d_part_bdot = nullptr;  // pass B spectral r . z partials (float32 lines)
  bool lean32 = false;            // every shape group on the register-FFT path: lean FP32 iteration
  bool dir32 = false;             // ... and on the staged passes: pass C forms p and steps x
  bool alpha64 = false;           // every shape group on the staged FP64 passes: pass A forms alpha
  int* d_chunk_range = nullptr;

  // workspaces
  Work<double> w64;
  Work<float> w32;
  bool w64_ready = false, w32_ready = false;
  double* d_part_mv = nullptr;
  double* d_part_a = nullptr;
  double* d_part_c = nullptr;
  double* d_part_b = nullptr;
  double* d_part_e = nullptr;
  double* d_ritz = nullptr;    // Ritz start: G, H (ncells*36 each) and coefficients (ncells*36)
  float* d_dinv32 = nullptr;   // Chebyshev smoothing: 1 / diag K(s) per voxel
  double* d_strain = nullptr;
  double* d_params = nullptr;  // per cell: [0, ncells) FP64 tolerance, [ncells, 2 ncells) inner FP32 tolerance
  int* d_iparams = nullptr;    // per cell: iteration cap of the FP64 PCG [0, ncells), of an FP32 inner PCG [ncells, 2 ncells)
  double* d_ir = nullptr;      // rn_true (ncells*M), bn (ncells*M), mean (ncells*M*3)
  int* d_outer_done = nullptr;
  int* d_maxed = nullptr;      // per cell: iteration budget exhausted
  int* d_restart = nullptr;    // per cell: the next FP32 inner solve restarts from p = z
  double inner_tol = 1e-2;     // loosest relative tolerance of an FP32 inner solve
  bool inner_restart = false;  // restart every FP32 inner PCG from z_0 instead of carrying the direction over
  int* h_pinned = nullptr;     // ncells * 4 ints
  double* h_pinned_d = nullptr;
  std::vector<void*> allocs;
  int64_t bytes = 0;

  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::map<GraphKey, hipGraph_t> graph_src;

  template <typename T>
  int alloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
      return NLFH_ENOMEM;
    }
    allocs.push_back(*p);
    bytes += n * sizeof(T);
    return NLFH_OK;
  }
  template <typename T>
  int upload(T** p, const std::vector<T>& v) {
    TRY(alloc(p, v.size()));
    if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NLFH_OK;
  }
  int64_t field_len() const { return (int64_t)M * 3 * tot_vox; }
  // values of a caller's field array with m right-hand sides
  int64_t field_len_m(int m) const { return (int64_t)(ncells == 1 ? m : M) * 3 * tot_vox; }
  int64_t spec_len() const { return (int64_t)M * 3 * tot_spec; }
};

namespace {

// Real symmetric symbol of K(1) from the 27-point stencil, inverted with the reference's
// regularisation (nlfiber/linear.py:154-165): Kh(k) = sum_o S_o exp(-2 pi i k.o / n); Kh(0) = I;
// Kinv = inv(Kh + 1e-12 |tr Kh| I); Kinv(0) = 0. Stored as (k00, k11, k22, k12, k02, k01).
// dense Ke from B matrices at the Gauss points (the reference's construction, nlfiber/linear.py:57-71,108-112)
void element_matrix(double lam, double mu, double (&Ke)[24][24]) {
  double gp[2] = {0.5 - 1.0 / (2.0 * std::sqrt(3.0)), 0.5 + 1.0 / (2.0 * std::sqrt(3.0))};
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) Ke[i][j] = 0.0;
  double D[6][6] = {};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i][j] = lam;
  for (int i = 0; i < 3; ++i) D[i][i] = lam + 2 * mu;
  for (int i = 3; i < 6; ++i) D[i][i] = mu;
  for (int g = 0; g < 8; ++g) {
    const double x = gp[g & 1], y = gp[(g >> 1) & 1], z = gp[(g >> 2) & 1];
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
      const int ia = a & 1, ja = (a >> 1) & 1, ka = (a >> 2) & 1;
      const double fx = ia ? x : 1 - x, fy = ja ? y : 1 - y, fz = ka ? z : 1 - z;
      const double sx = ia ? 1 : -1, sy = ja ? 1 : -1, sz = ka ? 1 : -1;
      const double dx = sx * fy * fz, dy = fx * sy * fz, dz = fx * fy * sz;
      const int c = 3 * a;
      B[0][c] = dx; B[1][c + 1] = dy; B[2][c + 2] = dz;
      B[3][c + 1] = dz; B[3][c + 2] = dy;
      B[4][c] = dz; B[4][c + 2] = dx;
      B[5][c] = dy; B[5][c + 1] = dx;
    }
    double DB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 24; ++j) {
        double t = 0;
        for (int k = 0; k < 6; ++k) t += D[i][k] * B[k][j];
        DB[i][j] = t;
      }
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double t = 0;
        for (int k = 0; k < 6; ++k) t += B[k][i] * DB[k][j];
        Ke[i][j] += t / 8.0;
      }
  }
}

// An upper bound of the spectrum of D^-1 K(s), D = diag K(s), for every stiffness field s > 0: with
// K(s) = sum_e s_e P_e^T Ke P_e and every diagonal entry of Ke equal to Ke_00, u^T D u = sum_e s_e Ke_00 |P_e u|^2,
// so the Rayleigh quotient u^T K u / u^T D u is an s-weighted mean of element quotients, each at most
// lambda_max(Ke) / Ke_00 (5.318 at nu = 0.3). lambda_max by power iteration on the 24 x 24 PSD matrix.
double jacobi_bound(double lam, double mu) {
  double Ke[24][24];
  element_matrix(lam, mu, Ke);
  double v[24], w[24], lmax = 0.0;
  for (int i = 0; i < 24; ++i) v[i] = 1.0 + 0.01 * i * (i % 3 + 1) * ((i & 4) ? -1.0 : 1.0);
  for (int it = 0; it < 2000; ++it) {
    double nw = 0.0;
    for (int i = 0; i < 24; ++i) {
      double t = 0.0;
      for (int j = 0; j < 24; ++j) t += Ke[i][j] * v[j];
      w[i] = t;
      nw += t * t;
    }
    nw = std::sqrt(nw);
    double nv = 0.0;
    for (int i = 0; i < 24; ++i) nv += v[i] * v[i];
    lmax = nw / std::sqrt(nv);
    for (int i = 0; i < 24; ++i) v[i] = w[i] / nw;
  }
  return lmax / Ke[0][0];
}

void build_symbol(const Shape& sh, double E, double nu, std::vector<double>& out) {
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  double Ke[24][24];
  element_matrix(lam, mu, Ke);
  // stencil: S[o][a][c] = response at node offset o (in {-1,0,1}^3) of component a to a unit
  // displacement of component c at the origin = sum over elements containing both nodes
  double S[27][3][3] = {};
  for (int e = 0; e < 8; ++e) {  // element with the origin at local node e
    const int ex = -(e & 1), ey = -((e >> 1) & 1), ez = -((e >> 2) & 1);  // element origin
    for (int b = 0; b < 8; ++b) {  // responding node
      const int ox = ex + (b & 1), oy = ey + ((b >> 1) & 1), oz = ez + ((b >> 2) & 1);
      const int oi = ((ox + 1) * 3 + (oy + 1)) * 3 + (oz + 1);
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) S[oi][a][c] += Ke[3 * b + a][3 * e + c];
    }
  }
  const int nx = sh.nx, ny = sh.ny, nz = sh.nz, nzh = nz / 2 + 1;
  out.assign((size_t)nx * ny * nzh * 6, 0.0);
  const long double two_pi = 6.283185307179586476925286766559L;
  for (int kx = 0; kx < nx; ++kx)
    for (int ky = 0; ky < ny; ++ky)
      for (int kz = 0; kz < nzh; ++kz) {
        long double K[3][3] = {};
        for (int ox = -1; ox <= 1; ++ox)
          for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
              const int oi = ((ox + 1) * 3 + (oy + 1)) * 3 + (oz + 1);
              // rfftn of the column response r(x) = S_x: Kh(k) = sum_x r(x) e^{-i k.x}; the
              // stencil is even, so the imaginary part cancels and Kh = sum S_o cos(k.o)
              const long double ph = two_pi * ((long double)kx * ox / nx + (long double)ky * oy / ny +
                                               (long double)kz * oz / nz);
              const long double cs = cosl(ph);
              for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) K[a][c] += S[oi][a][c] * cs;
            }
        if (kx == 0 && ky == 0 && kz == 0) continue;  // Kinv(0) = 0
        const long double tr = fabsl(K[0][0] + K[1][1] + K[2][2]);
        for (int a = 0; a < 3; ++a) K[a][a] += 1e-12L * tr;
        // symmetric 3x3 inverse by cofactors
        const long double c00 = K[1][1] * K[2][2] - K[1][2] * K[2][1];
        const long double c01 = K[1][2] * K[2][0] - K[1][0] * K[2][2];
        const long double c02 = K[1][0] * K[2][1] - K[1][1] * K[2][0];
        const long double det = K[0][0] * c00 + K[0][1] * c01 + K[0][2] * c02;
        const long double i00 = c00 / det;
        const long double i11 = (K[0][0] * K[2][2] - K[0][2] * K[2][0]) / det;
        const long double i22 = (K[0][0] * K[1][1] - K[0][1] * K[1][0]) / det;
        const long double i12 = (K[0][2] * K[1][0] - K[0][0] * K[1][2]) / det;
        const long double i02 = (K[0][1] * K[1][2] - K[0][2] * K[1][1]) / det;
        const long double i01 = (K[0][2] * K[2][1] - K[0][1] * K[2][2]) / det;
        double* o = &out[(((size_t)kx * ny + ky) * nzh + kz) * 6];
        o[0] = (double)i00; o[1] = (double)i11; o[2] = (double)i22;
        o[3] = (double)i12; o[4] = (double)i02; o[5] = (double)i01;
      }
}

// default tiling: one z value per lane up to 64, rows so that a block has <= 512 threads
void default_tiling(Shape& sh) {
  int zpt = 1;
  while (sh.nz / zpt > 64 || sh.nz % zpt) {
    ++zpt;
    if (zpt > 4) break;
  }
  sh.zpt = zpt;
  int lz = sh.nz / zpt, lzs = 1;
  while (lzs < lz) lzs <<= 1;
  int rows = std::max(3, 512 / lzs);
  // three LDS buffers of (ty + 2) rows x 6 values x nz (float64) must fit
  const int lds_rows = (int)(kLdsMax / ((size_t)3 * 6 * sh.nz * sizeof(double)));
  rows = std::min(rows, lds_rows);
  // small tiles keep every CU busy: the autotuner's measurements on gfx950 pick ty = 2-4 and xc = 2 (nx <= 32),
  // 8 (nx <= 64) or 16 (larger) for cubic cells from 16^3 to 128^3
  const int ty_pref = sh.ny <= 16 ? 2 : 4;
  sh.ty = std::max(1, std::min(std::min(rows - 2, ty_pref), sh.ny));
  const int xc_pref = sh.nx <= 32 ? 2 : (sh.nx <= 64 ? 8 : 16);
  sh.xc = std::max(1, std::min(xc_pref, sh.nx));
}

int build_mv_items(nlfh_plan* p) {
  p->mv.clear();
  p->groups.clear();
  std::vector<std::vector<int>> cell_items(p->ncells);
  for (int s = 0; s < (int)p->shapes.size(); ++s) {
    const Shape& sh = p->shapes[s];
    MvGroup g{s, (int)p->mv.size(), 0, 0};
    for (int c = 0; c < p->ncells; ++c) {
      if (p->cell_shape[c] != s) continue;
      for (int y0 = 0; y0 < sh.ny; y0 += sh.ty)
        for (int x0 = 0; x0 < sh.nx; x0 += sh.xc) {
          MvItem it{c, y0, x0, std::min(sh.xc, sh.nx - x0), std::min(sh.ty, sh.ny - y0), 0};
          g.max_ty = std::max(g.max_ty, it.ty);
          cell_items[c].push_back((int)p->mv.size());
          p->mv.push_back(it);
        }
    }
    g.count = (int)p->mv.size() - g.begin;
    if (g.count) p->groups.push_back(g);
  }
  // items of one cell are contiguous because cells of a shape are grouped; record ranges
  std::vector<int> range(2 * p->ncells);
  for (int c = 0; c < p->ncells; ++c) {
    range[2 * c] = cell_items[c].front();
    range[2 * c + 1] = cell_items[c].back() + 1;
    if (range[2 * c + 1] - range[2 * c] != (int)cell_items[c].size()) {
      set_error("internal: matvec items of a cell are not contiguous");
      return NLFH_EINVAL;
    }
  }
  if (p->d_mv) { (void)hipFree(p->d_mv); p->allocs.erase(std::find(p->allocs.begin(), p->allocs.end(), (void*)p->d_mv)); }
  if (p->d_mv_range) { (void)hipFree(p->d_mv_range); p->allocs.erase(std::find(p->allocs.begin(), p->allocs.end(), (void*)p->d_mv_range)); }
  TRY(p->upload(&p->d_mv, p->mv));
  TRY(p->upload(&p->d_mv_range, range));
  return NLFH_OK;
}

template <typename T>
size_t line_len_for(int nx, int64_t npos);

// FP32 line items of pass B (per shape group: the fast path's positions per block, lpb32 or the default, or the
// generic length) with their per-cell ranges; `rebuild` frees the device tables of an earlier build
int build_lines32(nlfh_plan* p, bool rebuild) {
  p->lines32.clear();
  std::vector<int> lrange32(2 * p->ncells);
  for (PrecGroup& g : p->pgroups) {
    g.line_begin[1] = (int)p->lines32.size();
    for (int c = 0; c < p->ncells; ++c) {
      if (p->cell_shape[c] != g.shape) continue;
      const CellDesc& cd = p->cells[c];
      const int64_t npos = (int64_t)cd.ny * cd.nzh;
      const int len = g.fast ? (g.lpb32 > 0 ? g.lpb32 : fast_line_len(g.fast, 4)) : (int)line_len_for<float>(cd.nx, npos);
      lrange32[2 * c] = (int)p->lines32.size();
      for (int64_t q = 0; q < npos; q += len)
        p->lines32.push_back({c, (int)q, (int)std::min<int64_t>(len, npos - q), 0});
      lrange32[2 * c + 1] = (int)p->lines32.size();
    }
    g.line_count[1] = (int)p->lines32.size() - g.line_begin[1];
  }
  if (rebuild) {
    for (void** q : {(void**)&p->d_lines32, (void**)&p->d_line_range32}) {
      if (*q) {
        (void)hipFree(*q);
        p->allocs.erase(std::find(p->allocs.begin(), p->allocs.end(), *q));
        *q = nullptr;
      }
    }
  }
  TRY(p->upload(&p->d_lines32, p->lines32));
  TRY(p->upload(&p->d_line_range32, lrange32));
  return NLFH_OK;
}

template <typename T>
size_t line_len_for(int nx, int64_t npos) {
  const size_t budget = 48 * 1024;
  size_t len = budget / (3 * (size_t)nx * sizeof(cplx<T>));
  if (len < 1) len = 1;
  if ((int64_t)len > npos) len = (size_t)npos;
  return len;
}

template <typename T>
void plane_lds(nlfh_plan* p, int& rpb, size_t& lds) {
  rpb = 32;
  for (int c = 0; c < p->ncells; ++c) {
    if (p->pgroups[p->cell_shape[c]].fast) continue;
    const CellDesc& cd = p->cells[c];
    const size_t plane = (size_t)cd.ny * cd.nzh * sizeof(cplx<T>);
    const size_t row = (size_t)cd.nz * sizeof(cplx<T>);
    int fit = plane < kLdsMax ? (int)((kLdsMax - plane) / row) : 0;
    rpb = std::min(rpb, std::max(1, fit));
  }
  lds = 0;
  for (int c = 0; c < p->ncells; ++c) {
    if (p->pgroups[p->cell_shape[c]].fast) continue;
    const CellDesc& cd = p->cells[c];
    const int npairs = (cd.ny + 1) / 2;
    lds = std::max(lds, prec_plane_lds<T>(cd.ny, cd.nz, std::min(rpb, npairs)));
  }
}

template <typename T>
int ensure_work(nlfh_plan* p, Work<T>& w) {
  const int64_t F = p->field_len();
  TRY(p->alloc(&w.X, F));
  TRY(p->alloc(&w.R, F));
  TRY(p->alloc(&w.B, F));
  TRY(p->alloc(&w.P[0], F));
  TRY(p->alloc(&w.P[1], F));
  TRY(p->alloc(&w.Ap, F));
  TRY(p->alloc(&w.Z, F));
  TRY(p->alloc(&w.s, p->tot_vox));
  TRY(p->alloc(&w.spec, p->spec_len()));
  const int nk = p->ncells * p->M;
  TRY(p->alloc(&w.st, 6 * nk));
  TRY(p->alloc(&w.ist, 2 * p->ncells));
  w.cs.rz = w.st;
  w.cs.alpha = w.st + nk;
  w.cs.beta = w.st + 2 * nk;
  w.cs.bn = w.st + 3 * nk;
  w.cs.rn = w.st + 4 * nk;
  w.cs.it = w.ist;
  w.cs.done = w.ist + p->ncells;
  w.cs.skip = nullptr;
  return NLFH_OK;
}

int get_ready64(nlfh_plan* p) {
  if (!p->w64_ready) {
    TRY(ensure_work(p, p->w64));
    p->w64.cs.tol = p->d_params;
    p->w64.cs.maxit = p->d_iparams;
    p->w64_ready = true;
  }
  return NLFH_OK;
}

int get_ready32(nlfh_plan* p) {
  if (!p->w32_ready) {
    TRY(ensure_work(p, p->w32));
    if (p->dir32) {
      TRY(p->alloc(&p->w32.rspec, p->spec_len()));
      p->w32.cs.alpha_prev = p->w32.st + 5 * p->ncells * p->M;
    }
    p->w32.cs.tol = p->d_params + p->ncells;
    p->w32.cs.maxit = p->d_iparams + p->ncells;
    p->w32_ready = true;
  }
  return NLFH_OK;
}

template <typename T>
T* sym_ptr(nlfh_plan* p);
template <>
double* sym_ptr<double>(nlfh_plan* p) { return p->d_sym64; }
template <>
float* sym_ptr<float>(nlfh_plan* p) { return p->d_sym32; }
template <typename T>
const cplx<T>* tw_ptr(nlfh_plan* p);
template <>
const cplx<double>* tw_ptr<double>(nlfh_plan* p) { return p->d_tw64; }
template <>
const cplx<float>* tw_ptr<float>(nlfh_plan* p) { return p->d_tw32; }

// K u (+ load) for every shape group. u may be null (load only).
template <typename T>
int run_matvec(nlfh_plan* p, const T* u, const T* p_old, T* p_new, T* out, const T* s, const double* beta,
               double* dot_part, const int* done, int m, int strain, int pupdate, int dot, T scale, hipStream_t st,
               const double* s64 = nullptr) {
  if (!strain && scale != T(1)) {
    set_error("internal: the matvec scales its output only together with the strain term");
    return NLFH_EINVAL;
  }
  if (!strain && u == nullptr) {
    set_error("internal: a matvec without the strain term needs an input field");
    return NLFH_EINVAL;
  }
  for (const MvGroup& g : p->groups) {
    const Shape& sh = p->shapes[g.shape];
    MatvecArgs<T> a{};
    a.cells = p->d_cells;
    a.items = p->d_mv + g.begin;
    a.u = u;
    a.p_old = p_old;
    a.p_new = p_new;
    a.out = out;
    a.s = s;
    a.s64 = s64;
    a.strain = p->d_strain;
    a.beta = beta;
    a.dot_part = dot_part;
    a.done = done;
    a.m = m;
    a.nitems = (int)p->mv.size();
    a.item_base = g.begin;
    a.LZ = sh.nz / sh.zpt;
    int lzs = 1;
    while (lzs < a.LZ) lzs <<= 1;
    a.LZS = lzs;
    a.lam16 = (T)(p->lam / 16.0);
    a.mu16 = (T)(p->mu / 16.0);
    a.scale_out = scale;
    HIP_TRY(launch_matvec<T>(a, g.count, sh.zpt, g.max_ty, strain, pupdate, dot, 0, st));
  }
  return NLFH_OK;
}

// One Chebyshev-Jacobi step for K x = -f on every cell: out = x + c1 (x - x_prev) - c2 D^-1 (K x + f).
int run_cheb_step(nlfh_plan* p, const float* x, const float* x_prev, float* out, const float* s32, const float* dinv,
                  float c1, float c2, int m, hipStream_t st) {
  for (const MvGroup& g : p->groups) {
    const Shape& sh = p->shapes[g.shape];
    MatvecArgs<float> a{};
    a.cells = p->d_cells;
    a.items = p->d_mv + g.begin;
    a.u = x;
    a.out = out;
    a.s = s32;
    a.strain = p->d_strain;
    a.m = m;
    a.nitems = (int)p->mv.size();
    a.item_base = g.begin;
    a.LZ = sh.nz / sh.zpt;
    int lzs = 1;
    while (lzs < a.LZ) lzs <<= 1;
    a.LZS = lzs;
    a.lam16 = (float)(p->lam / 16.0);
    a.mu16 = (float)(p->mu / 16.0);
    a.scale_out = 1.0f;
    a.x_prev = x_prev;
    a.dinv = dinv;
    a.c1 = c1;
    a.c2 = c2;
    HIP_TRY(launch_matvec<float>(a, g.count, sh.zpt, g.max_ty, 1, 0, 0, 1, st));
  }
  return NLFH_OK;
}

template <typename T>
PrecArgs<T> prec_args(nlfh_plan* p, int m) {
  PrecArgs<T> a{};
  a.cells = p->d_cells;
  a.planes = p->d_planes;
  a.lines = sizeof(T) == 8 ? p->d_lines64 : p->d_lines32;
  a.ft.lens = p->d_lens;
  a.ft.tw = tw_ptr<T>(p);
  a.ft.perm = p->d_perm;
  a.ft.iperm = p->d_iperm;
  a.ft.shape_len = p->d_shape_len;
  a.sym = sym_ptr<T>(p);
  a.m = m;
  a.nplanes = (int)p->planes.size();
  a.nlines = (int)(sizeof(T) == 8 ? p->lines64.size() : p->lines32.size());
  a.rows_per_batch = sizeof(T) == 8 ? p->rpb64 : p->rpb32;
  a.part_bdot = p->d_part_bdot;
  return a;
}

template <typename T>
const T* symsoa_ptr(nlfh_plan* p);
template <>
const double* symsoa_ptr<double>(nlfh_plan* p) { return p->d_symsoa64; }
template <>
const float* symsoa_ptr<float>(nlfh_plan* p) { return p->d_symsoa32; }

template <typename T>
int run_prec(nlfh_plan* p, PrecArgs<T>& a, int update, int dot, hipStream_t st) {
  const size_t lp = sizeof(T) == 8 ? p->lds_plane64 : p->lds_plane32;
  const size_t ll = sizeof(T) == 8 ? p->lds_line64 : p->lds_line32;
  const int pr = sizeof(T) == 8 ? 0 : 1;
  for (const PrecGroup& g : p->pgroups) {
    if (g.plane_count == 0) continue;
    a.plane_base = g.plane_begin;
    a.plane_count = g.plane_count;
    a.line_base = g.line_begin[pr];
    if (g.fast) {
      const Shape& sh = p->shapes[g.shape];
      const cplx<T>* tw = tw_ptr<T>(p) + p->lens[sh.len[0]].tw_off;
      const T* symp = symsoa_ptr<T>(p) + p->symsoa_off[g.shape];
      HIP_TRY(launch_prec_fast<T>(g.fast, a, g.plane_count, g.line_count[pr], tw, symp, update, dot, g.vec,
                                  sizeof(T) == 4 ? g.lpb32 : 0, sizeof(T) == 4 ? g.ppa32 : 1, st));
    } else {
      // the generic passes update r and x together and take r . z in pass C (lean mode never gets here)
      if (dot >= 2 || update == 1 || update == 4 || a.rspec != nullptr) return NLFH_EUNSUPPORTED;
      HIP_TRY(launch_prec<T>(a, g.plane_count, g.line_count[pr], update != 0 ? 1 : 0, dot, lp, ll, st));
    }
  }
  return NLFH_OK;
}

VecArgs vec_args(nlfh_plan* p, int m, const int* done) {
  VecArgs va{};
  va.cells = p->d_cells;
  va.chunks = p->d_chunks;
  va.nchunks = (int)p->chunks.size();
  va.m = m;
  va.done = done;
  return va;
}

// One PCG iteration (nlfiber/linear.py:190-200) with the buffers of `w`; `cur` selects the old
// search direction.
// Lean iteration (FP32 inner solves when every shape group is on the register-FFT path): r . z comes
// from the spectra in pass B (Parseval), so pass C writes z without reading r. With the staged passes
// (dir32) the residual is kept as its (z, y) spectrum r-hat and the search direction is formed in
// pass C instead of the matvec: iteration k runs
//   matvec Ap_k = K p_k (+ p_k . Ap_k)  ->  pass A alpha_k, r-hat -= alpha_k FFT_zy(Ap_k), ||r||
//   ->  pass B (r-hat in, z-hat out)  ->  pass C p_{k+1} = z_{k+1} + beta_k p_k
//   (odd k also x += alpha_{k-1} p_{k-1} + alpha_k p_k)  ->  beta / convergence,
// so neither r nor z is stored in real space (two field passes less than the plain order), the matvec
// reads p alone, and x is read and written every second iteration; launch_x_flush_even adds the step a
// cell owes when it stops after an even iteration. Same PCG recurrences (the transform is linear); the
// norms and dots differ from the direct sums only by rounding.
template <typename T>
bool lean(const nlfh_plan* p) {
  return sizeof(T) == 4 && p->lean32;
}

template <typename T>
int cg_iteration(nlfh_plan* p, Work<T>& w, int m, int cur, hipStream_t st) {
  T* p_old = w.P[cur];
  T* p_new = w.P[1 - cur];
  const bool ln = lean<T>(p);
  const bool dir = ln && p->dir32;
  if (dir)  // p_old holds p_k
    TRY(run_matvec<T>(p, p_old, nullptr, nullptr, w.Ap, w.s, nullptr, p->d_part_mv, w.cs.done, m, 0, 0, 1, T(1), st));
  else
    TRY(run_matvec<T>(p, w.Z, p_old, p_new, w.Ap, w.s, w.cs.beta, p->d_part_mv, w.cs.done, m, 0, 1, 1, T(1), st));
  // FP64 with every group on the staged passes: pass A forms alpha from the matvec's partials (no alpha kernel)
  const bool own_alpha = !dir && sizeof(T) == 8 && p->alpha64;
  if (!dir && !own_alpha)
    HIP_TRY(launch_cg_alpha(w.cs, p->ncells, p->d_part_mv, (int)p->mv.size(), p->d_mv_range, st));
  PrecArgs<T> a = prec_args<T>(p, m);
  if (own_alpha) {
    a.part_mv = p->d_part_mv;
    a.mv_range = p->d_mv_range;
    a.nmv = (int)p->mv.size();
    a.rz = w.cs.rz;
    a.alpha_out = w.cs.alpha;
    a.alpha_prev_out = w.cs.alpha_prev;
  }
  a.r = w.R;
  a.r_out = w.R;
  a.Ap = w.Ap;
  a.x = w.X;
  a.p = p_new;
  a.alpha = w.cs.alpha;
  a.spec = w.spec;
  a.z = w.Z;
  a.part_a = p->d_part_a;
  a.part_c = p->d_part_c;
  a.done = w.cs.done;
  if (dir) {
    a.p = p_old;  // p_k: the iterate's step and the old direction of p_{k+1}
    a.z = p_new;  // receives p_{k+1}
    a.rspec = w.rspec;
    a.line_range = p->d_line_range32;
    a.rz = w.cs.rz;
    a.alpha_prev = w.cs.alpha_prev;
    a.part_mv = p->d_part_mv;  // pass A forms and stores alpha (no alpha kernel)
    a.mv_range = p->d_mv_range;
    a.nmv = (int)p->mv.size();
    a.alpha_out = w.cs.alpha;
    a.alpha_prev_out = w.cs.alpha_prev;
  }
  // cur == k & 1: every graph holds an even number of iterations
  TRY(run_prec<T>(p, a, dir ? 4 : 2, dir ? (cur ? 4 : 3) : (ln ? 2 : 1), st));
  if (ln) {
    HIP_TRY(launch_cg_beta(w.cs, p->ncells, p->d_part_a, (int)p->planes.size(), p->d_plane_range, p->d_part_bdot,
                           (int)p->lines32.size(), p->d_line_range32, st));
  } else {
    HIP_TRY(launch_cg_beta(w.cs, p->ncells, p->d_part_a, (int)p->planes.size(), p->d_plane_range, p->d_part_c,
                           (int)p->planes.size(), p->d_plane_range, st));
  }
  return NLFH_OK;
}

constexpr int kItersPerGraph = 16;  // even: the search-direction ping-pong returns to P[0]

template <typename T>
int get_graph(nlfh_plan* p, Work<T>& w, int m, hipGraphExec_t* out) {
  const GraphKey key{sizeof(T) == 8 ? 0 : 1, m, kItersPerGraph};
  auto itg = p->graphs.find(key);
  if (itg != p->graphs.end()) {
    *out = itg->second;
    return NLFH_OK;
  }
  hipGraph_t g;
  HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  int rc = NLFH_OK;
  for (int i = 0; i < kItersPerGraph && rc == NLFH_OK; ++i) rc = cg_iteration<T>(p, w, m, i & 1, p->stream);
  hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (rc != NLFH_OK) return rc;
  HIP_TRY(e);
  hipGraphExec_t ge;
  HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  p->graphs[key] = ge;
  p->graph_src[key] = g;
  *out = ge;
  return NLFH_OK;
}

// PCG on the mean-free space from the state prepared in w (R = b - K X, B holds b for ||b||):
// nlfiber/linear.py:173-203. Returns after every cell is done; iterations per cell in it_host.
// resume (FP32 direction iteration only): P[0] already holds the first search direction (the previous inner
// solve's last one, in this system's units), so the start computes r-hat_0 and r_0 . z_0 and stores z_0 aside
// instead of taking p_0 = z_0.
// restart_cells (device, nullable, with resume): cells flagged there take p_0 = z_0 after all.
template <typename T>
int run_cg(nlfh_plan* p, Work<T>& w, int m, const T* bvec, int* it_host, int use_graph, int maxit,
           bool resume = false, const int* restart_cells = nullptr) {
  hipStream_t st = p->stream;
  w.cs.m = m;  // per-(cell, rhs) scalars use stride m
  VecArgs va = vec_args(p, m, nullptr);
  HIP_TRY(launch_reduce<T>(va, bvec, nullptr, p->d_part_b, 0, st));
  PrecArgs<T> a = prec_args<T>(p, m);
  a.r = w.R;
  a.spec = w.spec;
  a.z = w.Z;
  a.part_a = p->d_part_a;
  a.part_c = p->d_part_c;
  a.done = w.cs.skip;
  const bool ln = lean<T>(p);
  const bool dir = ln && p->dir32;
  if (dir) {
    a.z = resume ? w.Z : w.P[0];  // p_0 = z_0 unless a direction is carried over
    a.rspec = w.rspec;  // r-hat_0
  }
  if (resume && !dir) {
    set_error("internal: only the FP32 direction iteration carries search directions over");
    return NLFH_EINVAL;
  }
  TRY(run_prec<T>(p, a, 0, ln ? 2 : 1, st));
  if (resume && restart_cells != nullptr)
    HIP_TRY(launch_copy_flagged<T>(vec_args(p, m, nullptr), restart_cells, w.Z, w.P[0], st));
  if (ln) {
    HIP_TRY(launch_cg_init(w.cs, p->ncells, p->d_part_a, (int)p->planes.size(), p->d_plane_range, p->d_part_bdot,
                           (int)p->lines32.size(), p->d_line_range32, p->d_part_b, (int)p->chunks.size(),
                           p->d_chunk_range, st));
  } else {
    HIP_TRY(launch_cg_init(w.cs, p->ncells, p->d_part_a, (int)p->planes.size(), p->d_plane_range, p->d_part_c,
                           (int)p->planes.size(), p->d_plane_range, p->d_part_b, (int)p->chunks.size(),
                           p->d_chunk_range, st));
  }
  if (!dir) HIP_TRY(launch_fill<T>(w.P[0], p->field_len(), T(0), st));  // p_{-1} = 0
  int* h_done = p->h_pinned;
  const int nc = p->ncells;
  int cur = 0;
  long launched = 0;
  for (;;) {
    HIP_TRY(hipMemcpyAsync(h_done, w.cs.done, nc * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    bool all = true;
    int ndone = 0;
    for (int c = 0; c < nc; ++c) {
      all = all && h_done[c];
      ndone += h_done[c] ? 1 : 0;
    }
    if (all) break;
    progress("%s PCG: %ld iterations launched, %d of %d cells converged", sizeof(T) == 8 ? "FP64" : "FP32",
             launched, ndone, nc);
    if (use_graph) {
      hipGraphExec_t ge;
      TRY(get_graph<T>(p, w, m, &ge));
      HIP_TRY(hipGraphLaunch(ge, st));
      launched += kItersPerGraph;
    } else {
      for (int i = 0; i < kItersPerGraph; ++i) {
        TRY(cg_iteration<T>(p, w, m, cur, st));
        cur ^= 1;
      }
      launched += kItersPerGraph;
    }
    if (launched > (long)maxit + 2 * kItersPerGraph) {
      set_error("internal: the solver did not flag completion within maxit iterations");
      return NLFH_EHIP;
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (dir) HIP_TRY(launch_x_flush_even<float>(vec_args(p, m, nullptr), w.X, w.P[0], w.cs.alpha, w.cs.it, st));
  }
  HIP_TRY(hipMemcpyAsync(it_host, w.cs.it, nc * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return NLFH_OK;
}

int enter(nlfh_plan* p, void* stream) {
  HIP_TRY(hipSetDevice(p->device));
  HIP_TRY(hipEventRecord(p->ev, (hipStream_t)stream));
  HIP_TRY(hipStreamWaitEvent(p->stream, p->ev, 0));
  return NLFH_OK;
}

int leave(nlfh_plan* p, void* stream) {
  HIP_TRY(hipEventRecord(p->ev, p->stream));
  HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, p->ev, 0));
  return NLFH_OK;
}

// Small parameters reach the device as kernel arguments (copied at launch), never as asynchronous
// copies from pageable host memory, which may be read after the caller's buffer is gone.
int set_strain(nlfh_plan* p, const double* Es_host, int m, hipStream_t st) {
  SmallParams v{};
  for (int i = 0; i < m * 6; ++i) v.d[i] = Es_host[i];
  HIP_TRY(launch_set_small(v, m * 6, p->d_strain, nullptr, st));
  return NLFH_OK;
}

int set_maxit(nlfh_plan* p, int maxit, hipStream_t st) {
  HIP_TRY(launch_fill<int>(p->d_iparams, 2 * (int64_t)p->ncells, maxit, st));
  return NLFH_OK;
}

int check_m(nlfh_plan* p, int m) {
  CHECK_ARG(m >= 1 && m <= kMaxRhs, "m must be in 1..6");
  CHECK_ARG(p->ncells == 1 || m == p->m_layout, "batched plans take m == m_layout right-hand sides");
  return NLFH_OK;
}

// mean projection of a FP64 field (nlfiber/linear.py:95-96): per (cell, rhs, component)
int project_mean(nlfh_plan* p, double* v, int m, hipStream_t st) {
  VecArgs va = vec_args(p, m, nullptr);
  HIP_TRY(launch_reduce<double>(va, v, nullptr, p->d_part_b, 1, st));
  double* mean = p->d_ir + 2 * p->ncells * p->M;  // [cell][rhs*3 + c], stride 3M
  HIP_TRY(launch_sum_partials(p->d_part_b, (int)p->chunks.size(), p->d_chunk_range, p->ncells, 3 * m, 1.0, p->d_cells,
                              3 * p->M, mean, st));
  va.m = p->M;  // the mean buffer uses the plan's rhs stride
  va.nchunks = (int)p->chunks.size();
  VecArgs vs = va;
  HIP_TRY(launch_sub_mean<double>(vs, v, mean, st));
  return NLFH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
extern "C" {

const char* nlfh_last_error(void) { return nlfh::g_err.c_str(); }

int nlfh_version(void) { return 1; }

int nlfh_plan_create(const int* shapes, int ncells, double E, double nu, int m_layout, nlfh_plan** out) {
  CHECK_ARG(out != nullptr, "out is null");
  *out = nullptr;
  CHECK_ARG(shapes != nullptr && ncells >= 1 && ncells <= NLFH_MAX_CELLS, "ncells must be in 1..4096");
  CHECK_ARG(m_layout >= 1 && m_layout <= kMaxRhs, "m_layout must be in 1..6");
  CHECK_ARG(nu > -1.0 && nu < 0.5 && E > 0.0, "E must be positive and nu in (-1, 0.5)");
  nlfh_plan* p = new nlfh_plan();
  auto fail = [&](int rc) {
    nlfh_plan_destroy(p);
    return rc;
  };
  HIP_TRY(hipGetDevice(&p->device));
  p->ncells = ncells;
  p->m_layout = m_layout;
  p->M = ncells == 1 ? kMaxRhs : m_layout;
  p->E = E;
  p->nu = nu;
  p->lam = E * nu / ((1 + nu) * (1 - 2 * nu));
  p->mu = E / (2 * (1 + nu));
  // distinct shapes in order of first appearance; cells grouped by shape in the tables
  std::vector<int> lens_n;
  for (int c = 0; c < ncells; ++c) {
    const int nx = shapes[3 * c], ny = shapes[3 * c + 1], nz = shapes[3 * c + 2];
    if (nx < 1 || ny < 1 || nz < 1 || nx > 512 || ny > 512 || nz > 256) {
      set_error("cell dimensions must be in 1..512 (x, y) and 1..256 (z)");
      return fail(NLFH_EUNSUPPORTED);
    }
    int sid = -1;
    for (int s = 0; s < (int)p->shapes.size(); ++s)
      if (p->shapes[s].nx == nx && p->shapes[s].ny == ny && p->shapes[s].nz == nz) sid = s;
    if (sid < 0) {
      Shape sh{nx, ny, nz, 0, 0, 0, 0, {0, 0, 0}};
      default_tiling(sh);
      if (sh.zpt > 4 || sh.nz % sh.zpt) {
        set_error("nz must be a multiple of z values per lane with at most 64 lanes (nz <= 256)");
        return fail(NLFH_EUNSUPPORTED);
      }
      const int dims[3] = {nx, ny, nz};
      for (int a = 0; a < 3; ++a) {
        std::vector<int> rad;
        if (!factor(dims[a], rad)) {
          set_error("every cell dimension must be a product of 2, 3 and 5 (FFT preconditioner)");
          return fail(NLFH_EUNSUPPORTED);
        }
        int li = -1;
        for (int l = 0; l < (int)lens_n.size(); ++l)
          if (lens_n[l] == dims[a]) li = l;
        if (li < 0) {
          li = (int)lens_n.size();
          lens_n.push_back(dims[a]);
        }
        sh.len[a] = li;
      }
      sid = (int)p->shapes.size();
      p->shapes.push_back(sh);
    }
    p->cell_shape.push_back(sid);
  }
  // cell order in memory = order given (the caller's arrays); offsets
  int64_t ov = 0, os = 0;
  for (int c = 0; c < ncells; ++c) {
    const Shape& sh = p->shapes[p->cell_shape[c]];
    CellDesc cd{};
    cd.nx = sh.nx; cd.ny = sh.ny; cd.nz = sh.nz; cd.nzh = sh.nz / 2 + 1;
    cd.shape_id = p->cell_shape[c];
    cd.nvox = (int64_t)sh.nx * sh.ny * sh.nz;
    cd.nspec = (int64_t)sh.nx * sh.ny * cd.nzh;
    cd.off_vox = ov;
    cd.off_field = (int64_t)p->M * 3 * ov;
    cd.off_spec = (int64_t)p->M * 3 * os;
    ov += cd.nvox;
    os += cd.nspec;
    p->cells.push_back(cd);
  }
  p->tot_vox = ov;
  p->tot_spec = os;
  // symbols
  std::vector<double> sym64;
  for (Shape& sh : p->shapes) {
    std::vector<double> s;
    build_symbol(sh, E, nu, s);
    sh.off_sym = (int64_t)sym64.size() / 6;
    sym64.insert(sym64.end(), s.begin(), s.end());
  }
  p->h_sym = sym64;
  for (int c = 0; c < ncells; ++c) p->cells[c].off_sym = p->shapes[p->cell_shape[c]].off_sym * 6;
  // scale by 1 / N of its shape for the device copies
  std::vector<double> symd(sym64.size());
  std::vector<float> symf(sym64.size());
  for (const Shape& sh : p->shapes) {
    const double inv = 1.0 / ((double)sh.nx * sh.ny * sh.nz);
    const size_t n = (size_t)sh.nx * sh.ny * (sh.nz / 2 + 1) * 6;
    for (size_t i = 0; i < n; ++i) {
      symd[sh.off_sym * 6 + i] = sym64[sh.off_sym * 6 + i] * inv;
      symf[sh.off_sym * 6 + i] = (float)symd[sh.off_sym * 6 + i];
    }
  }
  p->tot_sym = (int64_t)sym64.size();
  // FFT tables
  std::vector<cplx<double>> tw64;
  std::vector<cplx<float>> tw32;
  std::vector<int> perm, iperm;
  for (int n : lens_n) {
    FftLen fl{};
    fl.n = n;
    std::vector<int> rad;
    factor(n, rad);
    fl.nst = (int)rad.size();
    for (int i = 0; i < fl.nst; ++i) fl.rad[i] = rad[i];
    fl.tw_off = (int)tw64.size();
    fl.perm_off = (int)perm.size();
    fl.iperm_off = (int)iperm.size();
    for (int t = 0; t < n; ++t) {
      const long double ang = -6.283185307179586476925286766559L * t / n;
      const double re = (double)cosl(ang), im = (double)sinl(ang);
      tw64.push_back({re, im});
      tw32.push_back({(float)re, (float)im});
    }
    std::vector<int> idx = digit_reversal(n, rad);
    std::vector<int> inv(n);
    for (int q = 0; q < n; ++q) inv[idx[q]] = q;
    perm.insert(perm.end(), idx.begin(), idx.end());
    iperm.insert(iperm.end(), inv.begin(), inv.end());
    p->lens.push_back(fl);
  }
  std::vector<int> shape_len;
  for (const Shape& sh : p->shapes) {
    shape_len.push_back(sh.len[0]);
    shape_len.push_back(sh.len[1]);
    shape_len.push_back(sh.len[2]);
  }
  // plane and line items grouped by shape (a cell's items stay contiguous), chunk items per cell
  std::vector<int> prange(2 * ncells), crange;
  for (int sidx = 0; sidx < (int)p->shapes.size(); ++sidx) {
    const Shape& sh = p->shapes[sidx];
    PrecGroup g{};
    g.shape = sidx;
    g.fast = (sh.nx == sh.ny && sh.ny == sh.nz && fast_fft_size(sh.nx)) ? sh.nx : 0;
    g.plane_begin = (int)p->planes.size();
    g.line_begin[0] = (int)p->lines64.size();
    g.line_begin[1] = (int)p->lines32.size();
    for (int c = 0; c < ncells; ++c) {
      if (p->cell_shape[c] != sidx) continue;
      const CellDesc& cd = p->cells[c];
      prange[2 * c] = (int)p->planes.size();
      for (int comp = 0; comp < 3; ++comp)
        for (int x = 0; x < cd.nx; ++x) p->planes.push_back({c, comp, x, 0});
      prange[2 * c + 1] = (int)p->planes.size();
      const int64_t npos = (int64_t)cd.ny * cd.nzh;
      const int len = g.fast ? fast_line_len(g.fast, 8) : (int)line_len_for<double>(cd.nx, npos);
      for (int64_t q = 0; q < npos; q += len)
        p->lines64.push_back({c, (int)q, (int)std::min<int64_t>(len, npos - q), 0});
    }
    g.vec = 1;
    for (int c = 0; c < ncells; ++c)
      if (p->cell_shape[c] == sidx && (p->cells[c].off_field % 4 != 0 || p->cells[c].off_spec % 2 != 0)) g.vec = 0;
    g.plane_count = (int)p->planes.size() - g.plane_begin;
    g.line_count[0] = (int)p->lines64.size() - g.line_begin[0];
    p->pgroups.push_back(g);
  }
  for (int c = 0; c < ncells; ++c) {
    const CellDesc& cd = p->cells[c];
    crange.push_back((int)p->chunks.size());
    for (int64_t b = 0; b < cd.nvox; b += kChunk) p->chunks.push_back({c, 0, b, std::min<int64_t>(b + kChunk, cd.nvox)});
    crange.push_back((int)p->chunks.size());
  }
  // LDS of the generic passes (over the cells that use them)
  p->lds_line64 = p->lds_line32 = 0;
  for (int c = 0; c < ncells; ++c) {
    const PrecGroup& g = p->pgroups[p->cell_shape[c]];
    if (g.fast) continue;
    const CellDesc& cd = p->cells[c];
    const int64_t npos = (int64_t)cd.ny * cd.nzh;
    p->lds_line64 = std::max(p->lds_line64, (size_t)3 * cd.nx * line_len_for<double>(cd.nx, npos) * sizeof(cplx<double>));
    p->lds_line32 = std::max(p->lds_line32, (size_t)3 * cd.nx * line_len_for<float>(cd.nx, npos) * sizeof(cplx<float>));
  }
  plane_lds<double>(p, p->rpb64, p->lds_plane64);
  plane_lds<float>(p, p->rpb32, p->lds_plane32);
  if (p->lds_plane64 > kLdsMax || p->lds_plane32 > kLdsMax) {
    set_error("a y-z plane of the cell does not fit in LDS (ny * (nz/2+1) too large)");
    return fail(NLFH_EUNSUPPORTED);
  }
  // component-planar symbol copies of the fast shapes
  std::vector<double> soa64;
  std::vector<float> soa32;
  p->symsoa_off.assign(p->shapes.size(), -1);
  for (const PrecGroup& g : p->pgroups) {
    if (!g.fast) continue;
    const Shape& sh = p->shapes[g.shape];
    const int64_t nf = (int64_t)sh.nx * sh.ny * (sh.nz / 2 + 1);
    p->symsoa_off[g.shape] = (int64_t)soa64.size();
    const double inv = 1.0 / ((double)sh.nx * sh.ny * sh.nz);
    for (int e = 0; e < 6; ++e)
      for (int64_t f = 0; f < nf; ++f) {
        const double v = sym64[(sh.off_sym + f) * 6 + e] * inv;
        soa64.push_back(v);
        soa32.push_back((float)v);
      }
  }
  int rc = NLFH_OK;
  if ((rc = p->upload(&p->d_cells, p->cells))) return fail(rc);
  if ((rc = p->upload(&p->d_lens, p->lens))) return fail(rc);
  if ((rc = p->upload(&p->d_tw64, tw64))) return fail(rc);
  if ((rc = p->upload(&p->d_tw32, tw32))) return fail(rc);
  if ((rc = p->upload(&p->d_perm, perm))) return fail(rc);
  if ((rc = p->upload(&p->d_iperm, iperm))) return fail(rc);
  if ((rc = p->upload(&p->d_shape_len, shape_len))) return fail(rc);
  if ((rc = p->upload(&p->d_sym64, symd))) return fail(rc);
  if ((rc = p->upload(&p->d_sym32, symf))) return fail(rc);
  if ((rc = p->upload(&p->d_symsoa64, soa64))) return fail(rc);
  if ((rc = p->upload(&p->d_symsoa32, soa32))) return fail(rc);
  if ((rc = p->upload(&p->d_planes, p->planes))) return fail(rc);
  if ((rc = p->upload(&p->d_lines64, p->lines64))) return fail(rc);
  if ((rc = build_lines32(p, false))) return fail(rc);
  if ((rc = p->upload(&p->d_chunks, p->chunks))) return fail(rc);
  if ((rc = p->upload(&p->d_plane_range, prange))) return fail(rc);
  p->lean32 = true;
  p->dir32 = true;
  p->alpha64 = true;
  for (const PrecGroup& g : p->pgroups) {
    p->lean32 = p->lean32 && g.fast != 0;
    p->dir32 = p->dir32 && g.fast != 0 && fast_staged(g.fast, 4) && g.vec;
    p->alpha64 = p->alpha64 && g.fast != 0 && fast_staged(g.fast, 8) && g.vec;
  }
  if ((rc = p->upload(&p->d_chunk_range, crange))) return fail(rc);
  if ((rc = build_mv_items(p))) return fail(rc);
  // partials: the matvec item count may change with the tiling, so size for the worst case
  int64_t max_mv = 0;
  for (int c = 0; c < ncells; ++c) max_mv += (int64_t)p->cells[c].nx * p->cells[c].ny;
  const int M = p->M;
  if ((rc = p->alloc(&p->d_part_mv, (size_t)M * max_mv))) return fail(rc);
  if ((rc = p->alloc(&p->d_part_a, (size_t)M * p->planes.size()))) return fail(rc);
  if ((rc = p->alloc(&p->d_part_c, (size_t)M * p->planes.size()))) return fail(rc);
  // pass B partials: sized for the smallest positions per block the FP32 passes offer (half the default)
  int64_t max_lines32 = 0;
  for (int c = 0; c < ncells; ++c) {
    const PrecGroup& g = p->pgroups[p->cell_shape[c]];
    const int64_t npos = (int64_t)p->cells[c].ny * p->cells[c].nzh;
    const int64_t len = g.fast ? std::max(1, fast_line_len(g.fast, 4) / 2) : (int64_t)line_len_for<float>(p->cells[c].nx, npos);
    max_lines32 += (npos + len - 1) / len;
  }
  if ((rc = p->alloc(&p->d_part_bdot, (size_t)M * std::max<int64_t>(1, max_lines32)))) return fail(rc);
  if ((rc = p->alloc(&p->d_part_b, (size_t)3 * M * p->chunks.size()))) return fail(rc);
  if ((rc = p->alloc(&p->d_part_e, (size_t)36 * p->chunks.size()))) return fail(rc);
  if ((rc = p->alloc(&p->d_strain, (size_t)6 * kMaxRhs))) return fail(rc);
  if ((rc = p->alloc(&p->d_params, 2 * (size_t)ncells))) return fail(rc);
  if ((rc = p->alloc(&p->d_iparams, 2 * (size_t)ncells))) return fail(rc);
  if ((rc = p->alloc(&p->d_ir, (size_t)ncells * M * 5))) return fail(rc);
  if ((rc = p->alloc(&p->d_outer_done, ncells))) return fail(rc);
  if ((rc = p->alloc(&p->d_maxed, ncells))) return fail(rc);
  if ((rc = p->alloc(&p->d_restart, ncells))) return fail(rc);
  if (hipHostMalloc((void**)&p->h_pinned, sizeof(int) * 4 * ncells, hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    return fail(NLFH_ENOMEM);
  }
  if (hipHostMalloc((void**)&p->h_pinned_d, sizeof(double) * ncells * M * 4, hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    return fail(NLFH_ENOMEM);
  }
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev, hipEventDisableTiming) != hipSuccess) {
    set_error("stream/event creation failed");
    return fail(NLFH_EHIP);
  }
  *out = p;
  return NLFH_OK;
}

int nlfh_plan_destroy(nlfh_plan* p) {
  if (!p) return NLFH_OK;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : p->graph_src) (void)hipGraphDestroy(kv.second);
  for (void* a : p->allocs) (void)hipFree(a);
  if (p->h_pinned) (void)hipHostFree(p->h_pinned);
  if (p->h_pinned_d) (void)hipHostFree(p->h_pinned_d);
  if (p->ev) (void)hipEventDestroy(p->ev);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return NLFH_OK;
}

int nlfh_plan_set_tiling(nlfh_plan* p, int shape_index, int ty, int xc, int zpt) {
  CHECK_ARG(p != nullptr, "plan is null");
  CHECK_ARG(shape_index >= 0 && shape_index < (int)p->shapes.size(), "shape_index out of range");
  Shape& sh = p->shapes[shape_index];
  CHECK_ARG(zpt >= 1 && zpt <= 4 && sh.nz % zpt == 0 && sh.nz / zpt <= 64, "zpt must divide nz with nz/zpt <= 64");
  int lz = sh.nz / zpt, lzs = 1;
  while (lzs < lz) lzs <<= 1;
  CHECK_ARG(ty >= 1 && lzs * (std::min(ty, sh.ny) + 2) <= 512, "ty too large for a 512-thread block");
  CHECK_ARG(xc >= 1, "xc must be positive");
  const size_t lds = (size_t)3 * (std::min(ty, sh.ny) + 2) * 6 * sh.nz * sizeof(double);
  CHECK_ARG(lds <= kLdsMax, "tile exceeds LDS");
  sh.ty = std::min(ty, sh.ny);
  sh.xc = std::min(xc, sh.nx);
  sh.zpt = zpt;
  // cached graphs bake the launch tables: drop them
  HIP_TRY(hipStreamSynchronize(p->stream));
  for (auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : p->graph_src) (void)hipGraphDestroy(kv.second);
  p->graphs.clear();
  p->graph_src.clear();
  return build_mv_items(p);
}

int nlfh_plan_set_fft_tiling(nlfh_plan* p, int shape_index, int pass_b_positions, int pass_a_planes) {
  CHECK_ARG(p != nullptr, "plan is null");
  CHECK_ARG(shape_index >= 0 && shape_index < (int)p->shapes.size(), "shape_index out of range");
  PrecGroup* grp = nullptr;
  for (PrecGroup& g : p->pgroups)
    if (g.shape == shape_index) grp = &g;
  CHECK_ARG(grp != nullptr, "shape has no cells");
  CHECK_ARG(pass_a_planes == 1 || pass_a_planes == 2, "pass_a_planes must be 1 or 2");
  if (!grp->fast) {
    CHECK_ARG(pass_b_positions == 0 && pass_a_planes == 1, "the generic passes have no tunable block shape");
    return NLFH_OK;
  }
  CHECK_ARG(fast_pass_b_ok(grp->fast, pass_b_positions),
            "pass_b_positions must be 0, the default, or half or double the default");
  CHECK_ARG(pass_a_planes == 1 || fast_two_planes_ok(grp->fast),
            "pass_a_planes must be 1 for this cell side (its staged pass A holds one plane per block)");
  grp->ppa32 = pass_a_planes;  // every fast cell has an even nx: a block's two planes share cell and component
  grp->lpb32 = pass_b_positions == fast_line_len(grp->fast, 4) ? 0 : pass_b_positions;
  // cached graphs bake the line tables: drop them
  HIP_TRY(hipStreamSynchronize(p->stream));
  for (auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : p->graph_src) (void)hipGraphDestroy(kv.second);
  p->graphs.clear();
  p->graph_src.clear();
  return build_lines32(p, true);
}

int nlfh_plan_get_fft_tiling(const nlfh_plan* p, int shape_index, int* pass_b_positions, int* pass_b_default,
                             int* pass_a_planes) {
  CHECK_ARG(p != nullptr, "plan is null");
  CHECK_ARG(shape_index >= 0 && shape_index < (int)p->shapes.size(), "shape_index out of range");
  for (const PrecGroup& g : p->pgroups) {
    if (g.shape != shape_index) continue;
    const int def = g.fast ? fast_line_len(g.fast, 4) : 0;
    if (pass_b_positions) *pass_b_positions = g.lpb32 > 0 ? g.lpb32 : def;
    if (pass_b_default) *pass_b_default = def;
    if (pass_a_planes) *pass_a_planes = g.ppa32;
    return NLFH_OK;
  }
  set_error("shape has no cells");
  return NLFH_EINVAL;
}

int nlfh_plan_get_tiling(const nlfh_plan* p, int shape_index, int* ty, int* xc, int* zpt) {
  CHECK_ARG(p != nullptr, "plan is null");
  CHECK_ARG(shape_index >= 0 && shape_index < (int)p->shapes.size(), "shape_index out of range");
  const Shape& sh = p->shapes[shape_index];
  if (ty) *ty = sh.ty;
  if (xc) *xc = sh.xc;
  if (zpt) *zpt = sh.zpt;
  return NLFH_OK;
}

int nlfh_stiffness_scale(nlfh_plan* p, const double* rho, double emin, double pw, double* s, double* dsdrho,
                         void* stream) {
  CHECK_ARG(p && rho && (s || dsdrho), "null argument");
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  HIP_TRY(launch_stiffness_scale(rho, s, nullptr, dsdrho, p->tot_vox, emin, pw, st));
  return NLFH_OK;
}

int nlfh_apply_K(nlfh_plan* p, const double* s, const void* u, void* Ku, int m, int dtype, void* stream) {
  CHECK_ARG(p && s && u && Ku, "null argument");
  TRY(check_m(p, m));
  CHECK_ARG(dtype == NLFH_DTYPE_F64 || dtype == NLFH_DTYPE_F32, "dtype must be 0 (f64) or 1 (f32)");
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  if (dtype == NLFH_DTYPE_F64) {
    TRY(run_matvec<double>(p, (const double*)u, nullptr, nullptr, (double*)Ku, s, nullptr, nullptr, nullptr, m, 0, 0,
                           0, 1.0, st));
  } else {
    // the FP32 operator rounds s as it loads it (no conversion pass)
    TRY(run_matvec<float>(p, (const float*)u, nullptr, nullptr, (float*)Ku, nullptr, nullptr, nullptr, nullptr, m, 0,
                          0, 0, 1.0f, st, s));
  }
  return NLFH_OK;
}

int nlfh_load(nlfh_plan* p, const double* s, const double* Es, int m, double* f, void* stream) {
  CHECK_ARG(p && s && Es && f, "null argument");
  TRY(check_m(p, m));
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  TRY(set_strain(p, Es, m, st));
  TRY(run_matvec<double>(p, nullptr, nullptr, nullptr, f, s, nullptr, nullptr, nullptr, m, 1, 0, 0, 1.0, st));
  return NLFH_OK;
}

int nlfh_residual(nlfh_plan* p, const double* s, const double* Es, const double* u, int m, double* r, void* stream) {
  CHECK_ARG(p && s && Es && u && r, "null argument");
  TRY(check_m(p, m));
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  TRY(set_strain(p, Es, m, st));
  TRY(run_matvec<double>(p, u, nullptr, nullptr, r, s, nullptr, nullptr, nullptr, m, 1, 0, 0, 1.0, st));
  return NLFH_OK;
}

int nlfh_precond(nlfh_plan* p, const void* r, void* z, int m, int dtype, void* stream) {
  CHECK_ARG(p && r && z, "null argument");
  TRY(check_m(p, m));
  CHECK_ARG(dtype == NLFH_DTYPE_F64 || dtype == NLFH_DTYPE_F32, "dtype must be 0 (f64) or 1 (f32)");
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  if (dtype == NLFH_DTYPE_F64) {
    TRY(get_ready64(p));
    PrecArgs<double> a = prec_args<double>(p, m);
    a.r = (const double*)r;
    a.spec = p->w64.spec;
    a.z = (double*)z;
    TRY(run_prec<double>(p, a, 0, 0, st));
  } else {
    TRY(get_ready32(p));
    PrecArgs<float> a = prec_args<float>(p, m);
    a.r = (const float*)r;
    a.spec = p->w32.spec;
    a.z = (float*)z;
    TRY(run_prec<float>(p, a, 0, 0, st));
  }
  return NLFH_OK;
}

}  // extern "C"

namespace {

// Final per-(cell, rhs) relative residual of the FP64 solver: rn / bn (recursive residual).
int fetch_fp64_info(nlfh_plan* p, int m, double tol, int* it_out, int* conv_out, double* relres_out, const int* its) {
  const int nk = p->ncells * m;
  HIP_TRY(hipMemcpyAsync(p->h_pinned_d, p->w64.cs.rn, nk * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipMemcpyAsync(p->h_pinned_d + nk, p->w64.cs.bn, nk * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  for (int c = 0; c < p->ncells; ++c) {
    bool conv = true;
    for (int t = 0; t < m; ++t) {
      const double rr = p->h_pinned_d[c * m + t] / p->h_pinned_d[nk + c * m + t];
      if (relres_out) relres_out[c * m + t] = rr;
      conv = conv && (p->h_pinned_d[c * m + t] <= tol * p->h_pinned_d[nk + c * m + t]);
    }
    if (conv_out) conv_out[c] = conv ? 1 : 0;
    if (it_out) it_out[c] = its[c];
  }
  return NLFH_OK;
}

int solve_fp64(nlfh_plan* p, int m, const double* u0, double tol, int maxit, int* it_out, int* conv_out,
               double* relres_out, bool rhs_given) {
  hipStream_t st = p->stream;
  Work<double>& w = p->w64;
  HIP_TRY(launch_fill<double>(p->d_params, p->ncells, tol, st));
  TRY(set_maxit(p, maxit, st));
  const int64_t F = p->field_len();
  if (!rhs_given) {
    // b = -load(s, Es)  (nlfiber/linear.py:205-206)
    TRY(run_matvec<double>(p, nullptr, nullptr, nullptr, w.B, w.s, nullptr, nullptr, nullptr, m, 1, 0, 0, -1.0, st));
  }
  if (u0) {
    HIP_TRY(launch_copy_d(u0, w.X, p->field_len_m(m), st));
    TRY(project_mean(p, w.X, m, st));
    if (rhs_given) {
      set_error("internal: rhs solves start from zero");
      return NLFH_EINVAL;
    }
    // r = b - K x = -(K x + load)
    TRY(run_matvec<double>(p, w.X, nullptr, nullptr, w.R, w.s, nullptr, nullptr, nullptr, m, 1, 0, 0, -1.0, st));
  } else {
    HIP_TRY(launch_fill<double>(w.X, F, 0.0, st));
    HIP_TRY(launch_copy_d(w.B, w.R, F, st));
  }
  std::vector<int> its(p->ncells, 0);
  TRY(run_cg<double>(p, w, m, w.B, its.data(), 1, maxit));
  return fetch_fp64_info(p, m, tol, it_out, conv_out, relres_out, its.data());
}


// FP32 PCG steps inside FP64 iterative refinement: x += K^-1 r solved in FP32 from the scaled true
// residual r = b - K x (FP64), until ||b - K x|| <= tol ||b|| per right-hand side (true FP64
// residual). Scalars of the inner PCG (dots, alpha, beta) are FP64.
// With the FP32 direction iteration the inner solves do not restart: each refinement step replaces the
// recursive residual by the true one and the next inner PCG continues along the last search direction of the
// previous one (rescaled to the new right-hand side r / ||r||; its first alpha is r.z / p.Ap with the new
// residual). A restart from p = z would discard the Krylov information gathered so far and cost a few
// percent more iterations per refinement step (residual replacement: van der Vorst and Ye, SIAM J. Sci.
// Comput. 22 (2000); reliable updates in mixed-precision CG: Clark et al., Comput. Phys. Commun. 181 (2010)).
int solve_mixed(nlfh_plan* p, int m, const double* u0, double tol, int maxit, int* it_out, int* conv_out,
                double* relres_out) {
  hipStream_t st = p->stream;
  TRY(get_ready32(p));
  Work<double>& w = p->w64;
  Work<float>& v = p->w32;
  const int nc = p->ncells;
  const int nk = nc * m;  // per-(cell, rhs) arrays use stride m
  double* rn_true = p->d_ir;
  double* bn = p->d_ir + nk;
  HIP_TRY(launch_fill<double>(p->d_params, 2 * (int64_t)nc, tol, st));
  TRY(set_maxit(p, maxit, st));
  HIP_TRY(launch_d2f(w.s, v.s, p->tot_vox, st));
  TRY(run_matvec<double>(p, nullptr, nullptr, nullptr, w.B, w.s, nullptr, nullptr, nullptr, m, 1, 0, 0, -1.0, st));
  VecArgs va = vec_args(p, m, nullptr);
  HIP_TRY(launch_reduce<double>(va, w.B, nullptr, p->d_part_b, 0, st));
  HIP_TRY(launch_norms(p->d_part_b, (int)p->chunks.size(), p->d_chunk_range, nc, m, m, 1e-300, bn, nullptr, 0.0,
                       nullptr, nullptr, nullptr, 0.0, 0.0, st));
  if (u0) {
    HIP_TRY(launch_copy_d(u0, w.X, p->field_len_m(m), st));
    TRY(project_mean(p, w.X, m, st));
  } else {
    HIP_TRY(launch_fill<double>(w.X, p->field_len(), 0.0, st));
  }
  HIP_TRY(launch_fill<int>(p->d_maxed, nc, 0, st));
  v.cs.skip = p->d_outer_done;
  std::vector<int> total(nc, 0), its(nc, 0), stall(nc, 0);
  std::vector<double> rn_prev((size_t)nk, 0.0);
  const bool keep = p->dir32 && !p->inner_restart;
  // per cell: the last inner solve continued a carried direction; the next one restarts; the cell restarts for the
  // rest of the solve (a continued solve hit its iteration cap)
  std::vector<int> resumed(nc, 0), restart_next(nc, 0), sticky(nc, 0), cap(nc, maxit);
  const double* h_rin = p->h_pinned_d + 2 * nk;  // last inner solve: recursive ||r||, then ||b|| (= ||r / ||r||||)
  int* h = p->h_pinned;
  for (int outer = 0; outer < 200; ++outer) {
    // true residual r = b - K x = -(K x + load), its norms and the outer convergence flags
    if (outer == 0 && !u0) HIP_TRY(launch_copy_d(w.B, w.R, p->field_len(), st));
    else TRY(run_matvec<double>(p, w.X, nullptr, nullptr, w.R, w.s, nullptr, nullptr, nullptr, m, 1, 0, 0, -1.0, st));
    HIP_TRY(launch_reduce<double>(va, w.R, nullptr, p->d_part_b, 0, st));
    HIP_TRY(launch_norms(p->d_part_b, (int)p->chunks.size(), p->d_chunk_range, nc, m, m, 0.0, rn_true, bn, tol,
                         p->d_outer_done, p->d_maxed, p->d_params + nc, p->inner_tol, 0.5, st));
    HIP_TRY(hipMemcpyAsync(h, p->d_outer_done, nc * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(p->h_pinned_d, rn_true, nk * sizeof(double), hipMemcpyDeviceToHost, st));
    if (keep && outer > 0) {
      HIP_TRY(hipMemcpyAsync(p->h_pinned_d + 2 * nk, v.cs.rn, nk * sizeof(double), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(p->h_pinned_d + 3 * nk, v.cs.bn, nk * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<int> odone(h, h + nc);  // h is reused by run_cg
    // stall: a refinement step that did not reduce some right-hand side's true residual by 10%
    // (FP32 inner solves that stopped helping) ends the cell unconverged; after an inner solve that continued a
    // carried direction the cell first gets one more step, restarted from p = z
    // drift: a true residual above 1.25x what the inner recursion reached (||r_prev|| times its recursive
    // relative residual) and above twice the inner tolerance's reduction means the FP32 recursion has drifted
    // from the true residual; the carried direction then belongs to the drifted problem and the next inner
    // solve restarts (right-hand sides the recursion took far below the tolerance may sit at their FP32 floor)
    bool stall_any = false;
    for (int c = 0; c < nc; ++c) {
      bool stalled = false, drift = false;
      for (int t = 0; t < m; ++t) {
        const double rn = p->h_pinned_d[c * m + t];
        if (outer > 0 && !odone[c] && !(rn < 0.9 * rn_prev[c * m + t])) stalled = true;
        if (keep && outer > 0 && !odone[c]) {
          const double pred = rn_prev[c * m + t] * h_rin[c * m + t] / std::max(h_rin[nk + c * m + t], 1e-300);
          if (!(rn <= std::max(1.25 * pred, 2.0 * p->inner_tol * rn_prev[c * m + t]))) drift = true;
        }
        rn_prev[c * m + t] = rn;
      }
      restart_next[c] = drift ? 1 : 0;
      if (stalled && !odone[c] && resumed[c]) {
        stalled = false;
        restart_next[c] = 1;
      }
      if (stalled && !odone[c]) {
        stall[c] = 1;
        odone[c] = 1;
        stall_any = true;
      }
    }
    if (stall_any) {
      for (int c = 0; c < nc; ++c) h[c] = odone[c];
      HIP_TRY(hipMemcpyAsync(p->d_outer_done, h, nc * sizeof(int), hipMemcpyHostToDevice, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    bool all = true;
    int ndone = 0;
    for (int c = 0; c < nc; ++c) {
      all = all && odone[c];
      ndone += odone[c] ? 1 : 0;
    }
    if (all) break;
    progress("refinement step %d: %d of %d cells done", outer, ndone, nc);
    // inner FP32 solve of K d = r / ||r|| from d = 0 (continuing the previous direction after the first step)
    bool resume = false, any_restart = false;
    for (int c = 0; c < nc; ++c) {
      resumed[c] = keep && outer > 0 && !odone[c] && !restart_next[c] && !sticky[c];
      resume = resume || resumed[c];
      any_restart = any_restart || (!odone[c] && !resumed[c]);
    }
    if (resume && any_restart) {
      int* hr = p->h_pinned + 2 * nc;  // pinned: the copy may run after the host moves on
      for (int c = 0; c < nc; ++c) hr[c] = resumed[c] ? 0 : 1;
      HIP_TRY(hipMemcpyAsync(p->d_restart, hr, nc * sizeof(int), hipMemcpyHostToDevice, st));
    }
    // a continued inner solve gets at most twice the cell's inner iterations so far (at least 256), plus 64: a
    // continuation that fails to converge is cut short, and the cell restarts from then on
    if (keep && outer > 0) {
      int* hc = p->h_pinned + 3 * nc;
      for (int c = 0; c < nc; ++c) {
        cap[c] = resumed[c] ? std::min(maxit, std::max(2 * total[c], 256) + 64) : maxit;
        hc[c] = cap[c];
      }
      HIP_TRY(hipMemcpyAsync(p->d_iparams + nc, hc, nc * sizeof(int), hipMemcpyHostToDevice, st));
    }
    progress("refinement step %d: %d inner solves continue their direction", outer,
             (int)std::count(resumed.begin(), resumed.end(), 1));
    HIP_TRY(launch_ir_down(vec_args(p, m, p->d_outer_done), w.R, v.R, v.X, rn_true, resume ? v.P[0] : nullptr, st));
    TRY(run_cg<float>(p, v, m, v.R, its.data(), 1, maxit, resume, any_restart ? p->d_restart : nullptr));
    HIP_TRY(launch_ir_acc(vec_args(p, m, p->d_outer_done), w.X, v.X, rn_true, keep ? v.P[0] : nullptr, v.P[1],
                          v.cs.it, st));
    int* maxed = p->h_pinned + nc;  // pinned: the copy below may run after this loop iteration
    bool any_max = false;
    for (int c = 0; c < nc; ++c) {
      if (!odone[c] && resumed[c] && its[c] >= cap[c]) sticky[c] = 1;
      if (!odone[c]) total[c] += its[c];
      maxed[c] = total[c] >= maxit ? 1 : 0;
      any_max = any_max || maxed[c];
    }
    if (any_max) {
      HIP_TRY(hipMemcpyAsync(p->d_maxed, maxed, nc * sizeof(int), hipMemcpyHostToDevice, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
  }
  // outputs: true relative residual per rhs
  HIP_TRY(hipMemcpyAsync(p->h_pinned_d, rn_true, nk * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(p->h_pinned_d + nk, bn, nk * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int c = 0; c < nc; ++c) {
    bool conv = true;
    for (int t = 0; t < m; ++t) {
      const double rr = p->h_pinned_d[c * m + t] / p->h_pinned_d[nk + c * m + t];
      if (relres_out) relres_out[c * m + t] = rr;
      conv = conv && (rr <= tol);
    }
    if (conv_out) conv_out[c] = conv ? 1 : 0;
    if (it_out) it_out[c] = total[c];
  }
  return NLFH_OK;
}

}  // namespace

extern "C" {

int nlfh_solve(nlfh_plan* p, const double* s, const double* Es, int m, const double* u0, double* u_out, double tol,
               int maxit, int precision, int* it_out, int* converged_out, double* relres_out, void* stream) {
  CHECK_ARG(p && s && Es && u_out, "null argument");
  TRY(check_m(p, m));
  CHECK_ARG(tol >= 0.0, "tol must be >= 0");
  CHECK_ARG(precision == NLFH_PREC_FP64 || precision == NLFH_PREC_MIXED, "precision must be 0 or 1");
  TRY(enter(p, stream));
  TRY(get_ready64(p));
  TRY(set_strain(p, Es, m, p->stream));
  HIP_TRY(launch_copy_d(s, p->w64.s, p->tot_vox, p->stream));
  if (precision == NLFH_PREC_FP64) {
    TRY(solve_fp64(p, m, u0, tol, maxit, it_out, converged_out, relres_out, false));
  } else {
    TRY(solve_mixed(p, m, u0, tol, maxit, it_out, converged_out, relres_out));
  }
  HIP_TRY(launch_copy_d(p->w64.X, u_out, p->field_len_m(m), p->stream));
  return leave(p, stream);
}

int nlfh_solve_rhs(nlfh_plan* p, const double* s, const double* b, int m, double* x_out, double tol, int maxit,
                   int precision, int* it_out, int* converged_out, double* relres_out, void* stream) {
  CHECK_ARG(p && s && b && x_out, "null argument");
  TRY(check_m(p, m));
  CHECK_ARG(precision == NLFH_PREC_FP64, "solve_rhs supports precision 0");
  TRY(enter(p, stream));
  TRY(get_ready64(p));
  HIP_TRY(launch_copy_d(s, p->w64.s, p->tot_vox, p->stream));
  HIP_TRY(launch_copy_d(b, p->w64.B, p->field_len_m(m), p->stream));
  // the reference projects b (nlfiber/linear.py:176)
  TRY(project_mean(p, p->w64.B, m, p->stream));
  TRY(solve_fp64(p, m, nullptr, tol, maxit, it_out, converged_out, relres_out, true));
  HIP_TRY(launch_copy_d(p->w64.X, x_out, p->field_len_m(m), p->stream));
  return leave(p, stream);
}

int nlfh_bilinear(nlfh_plan* p, const double* s, const double* Es, const double* u, int m, double* C_out,
                  double* dCds_out, void* stream) {
  CHECK_ARG(p && s && Es && u && C_out, "null argument");
  TRY(check_m(p, m));
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  TRY(set_strain(p, Es, m, st));
  EnergyArgs a{};
  a.cells = p->d_cells;
  a.chunks = p->d_chunks;
  a.nchunks = (int)p->chunks.size();
  a.m = m;
  a.eb = std::max(1, 256 / m);
  a.u = u;
  a.s = s;
  a.strain = p->d_strain;
  a.part = p->d_part_e;
  a.dcds = dCds_out;
  a.lam16 = p->lam / 16.0;
  a.mu16 = p->mu / 16.0;
  HIP_TRY(launch_energy(a, st));
  // C = sum / N per cell: rows = m*m; divide by N per cell (sum_partials takes one scale: loop cells)
  HIP_TRY(launch_sum_partials(p->d_part_e, (int)p->chunks.size(), p->d_chunk_range, p->ncells, m * m, 1.0, p->d_cells,
                              m * m, C_out, st));
  return NLFH_OK;
}

int nlfh_ritz_start(nlfh_plan* p, const double* s, const double* Es, int m, const double* u_in, double* u_out,
                    double* coef_out, void* stream) {
  CHECK_ARG(p && s && Es && u_in && u_out, "null argument");
  TRY(check_m(p, m));
  TRY(enter(p, stream));
  TRY(get_ready64(p));
  const int nc = p->ncells;
  if (!p->d_ritz) TRY(p->alloc(&p->d_ritz, (size_t)3 * 36 * nc));
  hipStream_t st = p->stream;
  Work<double>& w = p->w64;
  TRY(set_strain(p, Es, m, st));
  HIP_TRY(launch_copy_d(u_in, w.X, p->field_len_m(m), st));
  // K P and the load f = K_uE E; the correctors solve K u = -f (nlfiber/linear.py:212-216)
  TRY(run_matvec<double>(p, w.X, nullptr, nullptr, w.Ap, s, nullptr, nullptr, nullptr, m, 0, 0, 0, 1.0, st));
  TRY(run_matvec<double>(p, nullptr, nullptr, nullptr, w.B, s, nullptr, nullptr, nullptr, m, 1, 0, 0, 1.0, st));
  VecArgs va = vec_args(p, m, nullptr);
  const int nch = (int)p->chunks.size();
  double* G = p->d_ritz;
  double* H = p->d_ritz + (size_t)36 * nc;
  double* A = p->d_ritz + (size_t)72 * nc;
  HIP_TRY(launch_gram(va, w.X, w.Ap, p->d_part_e, st));
  HIP_TRY(launch_sum_partials(p->d_part_e, nch, p->d_chunk_range, nc, m * m, 1.0, nullptr, m * m, G, st));
  HIP_TRY(launch_gram(va, w.B, w.X, p->d_part_e, st));
  HIP_TRY(launch_sum_partials(p->d_part_e, nch, p->d_chunk_range, nc, m * m, 1.0, nullptr, m * m, H, st));
  std::vector<double> hg((size_t)2 * m * m * nc);
  HIP_TRY(hipMemcpyAsync(hg.data(), G, (size_t)m * m * nc * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hg.data() + (size_t)m * m * nc, H, (size_t)m * m * nc * sizeof(double),
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // per cell and right-hand side I: G a_I = -H[I][:], G_JK = P_J . K P_K, H_IJ = f_I . P_J
  std::vector<double> coef((size_t)m * m * nc, 0.0);
  for (int c = 0; c < nc; ++c) {
    const double* g = hg.data() + (size_t)c * m * m;
    const double* h = hg.data() + (size_t)m * m * nc + (size_t)c * m * m;
    // Cholesky of G with a relative diagonal floor; a field with no energy (P_J = 0) gets weight 0
    long double L[6][6] = {};
    double dmax = 0.0;
    for (int j = 0; j < m; ++j) dmax = std::max(dmax, g[j * m + j]);
    bool ok = std::isfinite(dmax) && dmax > 0.0;
    std::vector<int> act;
    for (int j = 0; ok && j < m; ++j)
      if (g[j * m + j] > 1e-14 * dmax) act.push_back(j);
    const int na = (int)act.size();
    for (int i = 0; ok && i < na; ++i) {
      for (int j = 0; j <= i; ++j) {
        long double v = g[act[i] * m + act[j]];
        for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
        if (i == j) {
          if (!(v > 1e-14L * g[act[i] * m + act[i]])) { ok = false; break; }
          L[i][i] = std::sqrt(v);
        } else {
          L[i][j] = v / L[j][j];
        }
      }
    }
    if (!ok) continue;  // zero coefficients: the zero start
    for (int I = 0; I < m; ++I) {
      long double y[6], x[6];
      for (int i = 0; i < na; ++i) {
        long double v = -h[I * m + act[i]];
        for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
        y[i] = v / L[i][i];
      }
      for (int i = na - 1; i >= 0; --i) {
        long double v = y[i];
        for (int k = i + 1; k < na; ++k) v -= L[k][i] * x[k];
        x[i] = v / L[i][i];
      }
      bool finite = true;
      for (int i = 0; i < na; ++i) finite = finite && std::isfinite((double)x[i]);
      if (!finite) continue;
      for (int i = 0; i < na; ++i) coef[(size_t)c * m * m + I * m + act[i]] = (double)x[i];
    }
  }
  HIP_TRY(hipMemcpyAsync(A, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(launch_combine(va, w.X, A, u_out, st));
  HIP_TRY(hipStreamSynchronize(st));  // coef (host) must outlive the copy
  if (coef_out) std::memcpy(coef_out, coef.data(), coef.size() * sizeof(double));
  return leave(p, stream);
}

double nlfh_jacobi_bound(double E, double nu) {
  if (!(E > 0.0) || !(nu > -1.0 && nu < 0.5)) return -1.0;
  return jacobi_bound(E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu)));
}

int nlfh_cheb_smooth(nlfh_plan* p, const double* s, const double* Es, int m, const double* u_in, double* u_out,
                     int degree, double lo, double hi, double* hi_used, void* stream) {
  CHECK_ARG(p && s && Es && u_in && u_out, "null argument");
  TRY(check_m(p, m));
  CHECK_ARG(degree >= 0 && degree <= 4096, "degree must be in 0..4096");
  const double bound = jacobi_bound(p->lam, p->mu);
  if (!(hi > 0.0)) hi = bound;
  CHECK_ARG(lo > 0.0 && lo < hi && std::isfinite(hi), "the interval must satisfy 0 < lo < hi");
  if (hi_used) *hi_used = hi;
  TRY(enter(p, stream));
  hipStream_t st = p->stream;
  const int64_t F = p->field_len_m(m);
  if (degree == 0) {
    if (u_out != u_in) HIP_TRY(launch_copy_d(u_in, u_out, F, st));
    return leave(p, stream);
  }
  TRY(get_ready32(p));
  if (!p->d_dinv32) TRY(p->alloc(&p->d_dinv32, p->tot_vox));
  Work<float>& w = p->w32;
  TRY(set_strain(p, Es, m, st));
  HIP_TRY(launch_d2f(s, w.s, p->tot_vox, st));
  HIP_TRY(launch_jacobi_dinv(p->d_cells, p->ncells, s, (p->lam + 4.0 * p->mu) / 9.0, p->d_dinv32, p->tot_vox, st));
  // three FP32 iterates in rotation: x_{k-1}, x_k, x_{k+1}
  float* buf[3] = {w.P[0], w.P[1], w.Ap};
  HIP_TRY(launch_d2f(u_in, buf[0], F, st));
  const double th = 0.5 * (hi + lo), de = 0.5 * (hi - lo), sig = th / de;
  double rho0 = 1.0 / sig;
  int prev = 0, cur = 0;
  for (int k = 0; k < degree; ++k) {
    double c1 = 0.0, c2 = 1.0 / th;
    if (k > 0) {
      const double rho1 = 1.0 / (2.0 * sig - rho0);
      c1 = rho1 * rho0;
      c2 = 2.0 * rho1 / de;
      rho0 = rho1;
    }
    const int nxt = (cur + 1) % 3 == prev ? (cur + 2) % 3 : (cur + 1) % 3;
    TRY(run_cheb_step(p, buf[cur], buf[prev], buf[nxt], w.s, p->d_dinv32, (float)c1, (float)c2, m, st));
    prev = cur;
    cur = nxt;
  }
  HIP_TRY(launch_f2d(buf[cur], u_out, F, st));
  return leave(p, stream);
}

int nlfh_plan_set_inner_tol(nlfh_plan* p, double inner_tol) {
  CHECK_ARG(p != nullptr, "plan is null");
  CHECK_ARG(inner_tol > 0.0 && inner_tol < 1.0, "inner_tol must be in (0, 1)");
  p->inner_tol = inner_tol;
  return NLFH_OK;
}

int nlfh_plan_set_inner_restart(nlfh_plan* p, int restart) {
  CHECK_ARG(p != nullptr, "plan is null");
  CHECK_ARG(restart == 0 || restart == 1, "restart must be 0 or 1");
  p->inner_restart = restart != 0;
  return NLFH_OK;
}

int nlfh_project_mean(nlfh_plan* p, double* v, int m, void* stream) {
  CHECK_ARG(p && v, "null argument");
  TRY(check_m(p, m));
  HIP_TRY(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;  // standalone operators run on the caller's stream
  TRY(project_mean(p, v, m, st));
  return NLFH_OK;
}

int nlfh_symbol(const nlfh_plan* p, int shape_index, double* kinv_out) {
  CHECK_ARG(p && kinv_out, "null argument");
  CHECK_ARG(shape_index >= 0 && shape_index < (int)p->shapes.size(), "shape_index out of range");
  const Shape& sh = p->shapes[shape_index];
  const size_t n = (size_t)sh.nx * sh.ny * (sh.nz / 2 + 1) * 6;
  std::memcpy(kinv_out, p->h_sym.data() + sh.off_sym * 6, n * sizeof(double));
  return NLFH_OK;
}

long long nlfh_plan_bytes(const nlfh_plan* p) { return p ? (long long)p->bytes : 0; }

}  // extern "C"
