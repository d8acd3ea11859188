// Kernel argument blocks and launchers shared by the library's translation units.
#pragma once
#include "nlfh_common.h"

namespace nlfh {

#define NLFH_FFT_MAX_STAGES 12

template <typename T>
struct MatvecArgs {
  const CellDesc* cells;
  const MvItem* items;
  const T* u;            // input field (nullptr: zero field, load only)
  const T* p_old;        // PUPDATE: previous search direction (u is z then)
  T* p_new;              // PUPDATE: new search direction p = z + beta p_old (owned nodes)
  T* out;                // output field
  const T* s;            // stiffness scale per voxel
  const double* s64;     // nullable: FP64 stiffness read (and rounded to T) instead of s
  const double* strain;  // STRAIN: macroscopic strains (m, 6) shared by all cells
  const double* beta;    // PUPDATE: beta per (cell, rhs)
  double* dot_part;      // DOT: partial sums of p . out per (rhs, item_base + item)
  const int* done;       // optional per-cell early-exit flags
  int m;
  int nitems;     // total matvec items of the plan (row stride of dot_part)
  int item_base;  // index of items[0] among all items
  int LZ;   // lanes holding a row of nz nodes (nz = LZ * ZPT)
  int LZS;  // lanes reserved per row: power of two >= LZ (lanes LZ..LZS-1 idle)
  T lam16, mu16;  // lambda / 16, mu / 16
  T scale_out;    // output multiplier (-1 gives b = -load)
  // CHEB (with STRAIN): one Chebyshev-Jacobi step, out = u + c1 (u - x_prev) - c2 dinv (K u + f)
  const T* x_prev;  // previous iterate (same layout as u)
  const T* dinv;    // 1 / diag K(s) per voxel (the diagonal is the same for the 3 components)
  T c1, c2;
};

struct FftLen {
  int n;
  int nst;
  int rad[NLFH_FFT_MAX_STAGES];
  int tw_off;    // offset of w_n^t, t in [0, n), in the twiddle tables
  int perm_off;  // offset of idx(p) (position p of a DIT input holds x[idx(p)]) in the perm table
  int iperm_off; // offset of the inverse map
};

template <typename T>
struct FftTables {
  const FftLen* lens;     // per distinct length
  const cplx<T>* tw;      // twiddles exp(-2 pi i t / n)
  const int* perm;
  const int* iperm;
  const int* shape_len;   // per shape id: index of (nx, ny, nz) in `lens` (3 ints)
};

template <typename T>
struct PrecArgs {
  const CellDesc* cells;
  const PlaneItem* planes;
  const LineItem* lines;
  FftTables<T> ft;
  const T* sym;          // per shape: [kx][ky][kz][6] real symmetric Kinv / (nx ny nz)
  const T* r;            // pass A input (and pass C for the r . z dot)
  T* r_out;              // pass A fused update: r - alpha Ap written here (may equal r)
  const T* Ap;           // pass A fused update (spectral update: the plane transformed instead of r)
  T* x;                  // pass A fused update x += alpha p; direction pass C: the iterate's paired step
  const T* p;            // pass A fused update; direction pass C: p_k, the old direction
  const double* alpha;   // per (cell, rhs)
  const double* alpha_prev;  // direction pass C, odd iterations: alpha of the previous iteration
  cplx<T>* spec;         // half spectra [cell][rhs][comp][x][ky][kz]
  cplx<T>* rspec;        // nullable: the residual's (z, y) spectrum, kept across iterations; when set, pass A
                         // writes it (not spec) and pass B reads it and writes spec
  T* z;                  // pass C output (direction pass C: p_{k+1}; odd k first reads p_{k-1} there)
  double* part_a;        // pass A: partial ||r||^2 per (rhs, plane item)
  double* part_c;        // pass C: partial r . z per (rhs, plane item)
  double* part_bdot;     // pass B (spectral dot): partial r . z per (rhs, line item)
  const int* line_range; // direction pass C: per cell [begin, end) of its line items (part_bdot columns)
  const double* rz;      // direction pass C: r . z of the previous iteration per (cell, rhs)
  // spectral pass A forms alpha itself from the matvec's p . Ap partials and stores it (and the previous
  // alpha) from one block per (cell, rhs)
  const double* part_mv; // p . Ap partials per (rhs, matvec item)
  const int* mv_range;   // per cell [begin, end) of its matvec items
  double* alpha_out;     // receives alpha per (cell, rhs)
  double* alpha_prev_out;  // receives the alpha it replaces
  int nmv;               // matvec items (row stride of part_mv)
  const int* done;       // per-cell early exit (nullable)
  int m;
  int nplanes;           // total plane items (row stride of part_a / part_c)
  int nlines;
  int rows_per_batch;    // row pairs per LDS batch in passes A and C
  int plane_base;        // first plane item of this launch
  int plane_count;       // plane items of this launch
  int line_base;         // first line item of this launch
};

struct VecArgs {
  const CellDesc* cells;
  const ChunkItem* chunks;
  int nchunks;
  int m;
  const int* done;  // per-cell early exit (nullable)
};

struct CgState {
  double* rz;     // [cell*m + rhs]
  double* alpha;
  double* alpha_prev;  // nullable: receives the previous alpha before cg_alpha overwrites it
  double* beta;
  double* bn;
  double* rn;
  int* it;        // [cell]
  int* done;      // [cell]
  const double* tol;  // [cell]: relative tolerance of each cell
  const int* maxit;   // [cell]: iteration cap
  const int* skip;    // nullable [cell]: cells that take no iterations (converged outer loop)
  int m;
};

struct EnergyArgs {
  const CellDesc* cells;
  const ChunkItem* chunks;
  int nchunks;
  int m;
  int eb;  // elements per block pass (blockDim = eb * m)
  int pad_;
  const double* u;       // correctors (SoA)
  const double* s;       // stiffness scale
  const double* strain;  // (m, 6)
  double* part;          // [chunk][m*m]
  double* dcds;          // nullable: [cell-major] (m, m, nvox) per cell at off_vox * m * m
  double lam16, mu16;
};

template <typename T>
hipError_t launch_matvec(const MatvecArgs<T>& a, int nblk, int zpt, int max_ty, int strain, int pupdate, int dot,
                         int cheb, hipStream_t st);
template <typename T>
size_t prec_plane_lds(int ny, int nz, int rows_per_batch);
template <typename T>
hipError_t launch_prec(const PrecArgs<T>& a, int nplanes, int nlines, int update, int dot, size_t lds_plane,
                       size_t lds_line, hipStream_t st);
bool fast_fft_size(int n);
bool fast_staged(int n, int elem_bytes);  // the staged passes A and C (the plane in LDS) exist for this length
int fast_line_len(int n, int elem_bytes);
// FP32 pass B of the fast path: lpb (ky, kz) positions per block is the default (0 or its value), or half or
// double of it
bool fast_pass_b_ok(int n, int lpb);
bool fast_two_planes_ok(int n);  // the FP32 staged pass A offers two x-planes per block at this length
// ppa: x-planes per block of the FP32 spectral pass A (1 or 2)
template <typename T>
hipError_t launch_prec_fast(int n, const PrecArgs<T>& a, int nplanes, int nlines, const cplx<T>* tw, const T* symp,
                            int update, int dot, int vec, int lpb, int ppa, hipStream_t st);
hipError_t launch_stiffness_scale(const double* rho, double* s64, float* s32, double* ds, int64_t n, double emin,
                                  double p, hipStream_t st);
hipError_t launch_d2f(const double* in, float* out, int64_t n, hipStream_t st);
hipError_t launch_copy_d(const double* in, double* out, int64_t n, hipStream_t st);
hipError_t launch_f2d(const float* in, double* out, int64_t n, hipStream_t st);
hipError_t launch_jacobi_dinv(const CellDesc* cells, int ncells, const double* s, double kdiag, float* dinv,
                              int64_t tot, hipStream_t st);
template <typename T>
hipError_t launch_fill(T* out, int64_t n, T v, hipStream_t st);
template <typename T>
hipError_t launch_reduce(const VecArgs& va, const T* a, const T* b, double* part, int mode, hipStream_t st);
// Parameters small enough to travel as kernel arguments (strains m x 6, iteration caps).
struct SmallParams {
  double d[36];
  int i[4];
};
template <typename T>
hipError_t launch_x_flush_even(const VecArgs& va, T* x, const T* p0, const double* alpha, const int* it,
                               hipStream_t st);
hipError_t launch_set_small(const SmallParams& v, int n, double* dd, int* di, hipStream_t st);
hipError_t launch_gram(const VecArgs& va, const double* a, const double* b, double* part, hipStream_t st);
hipError_t launch_combine(const VecArgs& va, const double* x, const double* coef, double* out, hipStream_t st);
hipError_t launch_sum_partials(const double* part, int nitems, const int* range, int ncells, int rows, double scale,
                               const CellDesc* cells, int out_stride, double* out, hipStream_t st);
template <typename T>
hipError_t launch_sub_mean(const VecArgs& va, T* v, const double* mean, hipStream_t st);
template <typename T>
hipError_t launch_copy_flagged(const VecArgs& va, const int* flag, const T* src, T* dst, hipStream_t st);
hipError_t launch_ir_down(const VecArgs& va, const double* r64, float* r32, float* d32, const double* scale,
                          float* pdir, hipStream_t st);
hipError_t launch_ir_acc(const VecArgs& va, double* x64, const float* d32, const double* scale, float* p0,
                         const float* p1, const int* it, hipStream_t st);
hipError_t launch_cg_init(const CgState& cs, int ncells, const double* pa, int np, const int* pr, const double* pz,
                          int nz, const int* zr, const double* pb, int nch, const int* cr, hipStream_t st);
hipError_t launch_cg_alpha(const CgState& cs, int ncells, const double* pmv, int nmv, const int* mr, hipStream_t st);
hipError_t launch_cg_beta(const CgState& cs, int ncells, const double* pa, int np, const int* pr, const double* pz,
                          int nz, const int* zr, hipStream_t st);
hipError_t launch_norms(const double* part, int nchunks, const int* crange, int ncells, int m, int M, double eps,
                        double* out, const double* bn, double tol, int* flag, const int* force, double* tol_in,
                        double tol_in_base, double tol_in_safety, hipStream_t st);
hipError_t launch_energy(const EnergyArgs& a, hipStream_t st);

}  // namespace nlfh
