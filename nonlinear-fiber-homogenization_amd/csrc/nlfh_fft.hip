// Reference-medium (Green's operator) preconditioner z = irfftn(Kinv . rfftn(r)), gfx950.
//
// Restates nlfiber/linear.py:167-170 with a hand-written mixed-radix (4, 2, 3, 5) FFT in LDS.
// Three passes per application, each one read and one write of the field:
//   A  per (cell, rhs, component, x-plane): real-to-complex along z (two real rows packed in one
//      complex transform), complex FFT along y, store the half spectrum plane [ky][kz];
//      optionally fused with the PCG update r -= alpha Ap, x += alpha p and ||r||^2;
//   B  per (cell, rhs, run of (ky, kz) positions): forward FFT along x for the three components,
//      z_a = sum_c Kinv_ac r_c with the real symmetric symbol (6 values per frequency, already
//      divided by nx ny nz, zero at k = 0), inverse FFT along x;
//   C  per (cell, rhs, component, x-plane): inverse FFT along y, complex-to-real along z (two rows
//      per transform), store z; optionally the partial r . z.
// The symbol of K(1) is real because the Q1 stencil is even (its imaginary part is checked to be
// at round-off against the reference construction in the tests), so storing its real part is
// exact.
//
// Transforms are in place: decimation in time (digit-reversed input, natural output) and its
// transpose, decimation in frequency (natural input, digit-reversed output), so no transform
// ever needs an in-place permutation.
#include "nlfh_kernels.h"

namespace nlfh {




// One radix-r DFT (r in 2, 3, 4, 5) of v[0..r-1], in place; sign = -1 forward, +1 inverse.
template <typename T>
__device__ __forceinline__ void small_dft(cplx<T>* v, int r, T sign) {
  if (r == 2) {
    const cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (r == 4) {
    const cplx<T> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
    const cplx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const cplx<T> t2 = cadd(a1, a3), t3 = csub(a1, a3);
    // multiply t3 by sign * i: forward (sign -1) gives -i
    const cplx<T> t3r = {-sign * t3.im, sign * t3.re};
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3r);
    v[3] = csub(t1, t3r);
  } else if (r == 3) {
    const T c1 = T(-0.5);
    const T s1 = sign * T(0.86602540378443864676);  // sin(2 pi / 3) with the transform's sign
    const cplx<T> a0 = v[0], a1 = v[1], a2 = v[2];
    const cplx<T> s = cadd(a1, a2), d = csub(a1, a2);
    const cplx<T> m = {a0.re + c1 * s.re, a0.im + c1 * s.im};
    // (i * s1) * d
    const cplx<T> id = {-s1 * d.im, s1 * d.re};
    v[0] = cadd(a0, s);
    v[1] = cadd(m, id);
    v[2] = csub(m, id);
  } else {  // r == 5
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = sign * T(0.95105651629515357212), s2 = sign * T(0.58778525229247312917);
    const cplx<T> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3], a4 = v[4];
    const cplx<T> b1 = cadd(a1, a4), b4 = csub(a1, a4);
    const cplx<T> b2 = cadd(a2, a3), b3 = csub(a2, a3);
    const cplx<T> m1 = {a0.re + c1 * b1.re + c2 * b2.re, a0.im + c1 * b1.im + c2 * b2.im};
    const cplx<T> m2 = {a0.re + c2 * b1.re + c1 * b2.re, a0.im + c2 * b1.im + c1 * b2.im};
    // i * (s1 b4 + s2 b3) and i * (s2 b4 - s1 b3)
    const cplx<T> n1 = {s1 * b4.re + s2 * b3.re, s1 * b4.im + s2 * b3.im};
    const cplx<T> n2 = {s2 * b4.re - s1 * b3.re, s2 * b4.im - s1 * b3.im};
    const cplx<T> in1 = {-n1.im, n1.re}, in2 = {-n2.im, n2.re};
    v[0] = cadd(a0, cadd(b1, b2));
    v[1] = cadd(m1, in1);
    v[4] = csub(m1, in1);
    v[2] = cadd(m2, in2);
    v[3] = csub(m2, in2);
  }
}

// In-place transforms of `nlines` lines in LDS: element p of line l at buf[l*lstride + p*estride].
// DIT: digit-reversed in, natural out. DIF: natural in, digit-reversed out. Threads of the block
// share the butterflies; `line_fast` orders work so that consecutive threads take consecutive lines.
template <typename T, bool DIF>
__device__ void fft_lds(cplx<T>* buf, const FftLen& fl, const cplx<T>* __restrict__ tw, int nlines,
                        int lstride, int estride, bool inverse, bool line_fast) {
  const int n = fl.n;
  if (n == 1) return;
  const T sign = inverse ? T(1) : T(-1);
  for (int si = 0; si < fl.nst; ++si) {
    const int s = DIF ? (fl.nst - 1 - si) : si;
    const int r = fl.rad[s];
    int m = 1;
    for (int t = 0; t < s; ++t) m *= fl.rad[t];
    const int L = m * r;
    const int tstep = n / L;
    const int nb = n / r;  // butterflies per line
    const int work = nb * nlines;
    for (int w = threadIdx.x; w < work; w += blockDim.x) {
      int line, b;
      if (line_fast) { line = w % nlines; b = w / nlines; }
      else { b = w % nb; line = w / nb; }
      const int blk = b / m, j = b - blk * m;
      cplx<T>* base = buf + (int64_t)line * lstride + (int64_t)(blk * L + j) * estride;
      cplx<T> v[5];
      for (int q = 0; q < r; ++q) v[q] = base[(int64_t)q * m * estride];
      if (!DIF) {
        for (int q = 1; q < r; ++q) {
          cplx<T> t = tw[(j * q * tstep) % n];
          if (inverse) t.im = -t.im;
          v[q] = cmul(v[q], t);
        }
      }
      small_dft<T>(v, r, sign);
      if (DIF) {
        for (int q = 1; q < r; ++q) {
          cplx<T> t = tw[(j * q * tstep) % n];
          if (inverse) t.im = -t.im;
          v[q] = cmul(v[q], t);
        }
      }
      for (int q = 0; q < r; ++q) base[(int64_t)q * m * estride] = v[q];
    }
    __syncthreads();
  }
}


// Pass A ---------------------------------------------------------------------------------------
template <typename T, bool UPDATE>
__global__ void __launch_bounds__(256) prec_pass_a(const PrecArgs<T> a) {
  extern __shared__ double smem_d[];
  __shared__ double red[4];
  const int item = a.plane_base + blockIdx.x;
  const PlaneItem pi = a.planes[item];
  if (a.done != nullptr && a.done[pi.cell]) return;
  const CellDesc cd = a.cells[pi.cell];
  const int rhs = blockIdx.y;
  const int ny = cd.ny, nz = cd.nz, nzh = cd.nzh;
  const int* sl = a.ft.shape_len + 3 * cd.shape_id;
  const FftLen fy = a.ft.lens[sl[1]];
  const FftLen fz = a.ft.lens[sl[2]];
  const cplx<T>* twy = a.ft.tw + fy.tw_off;
  const cplx<T>* twz = a.ft.tw + fz.tw_off;
  const int* ipz = a.ft.iperm + fz.iperm_off;
  const int* py = a.ft.perm + fy.perm_off;

  cplx<T>* plane = reinterpret_cast<cplx<T>*>(smem_d);  // [ny][nzh], natural y order
  cplx<T>* wsp = plane + (int64_t)ny * nzh;              // [rows_per_batch][nz]
  const int64_t nvox = cd.nvox;
  const int64_t fo = cd.off_field + ((int64_t)rhs * 3 + pi.comp) * nvox + (int64_t)pi.x * ny * nz;
  double acc = 0.0;
  T alpha = T(0);
  if (UPDATE) alpha = (T)a.alpha[pi.cell * a.m + rhs];

  const int npairs = (ny + 1) / 2;
  for (int pb = 0; pb < npairs; pb += a.rows_per_batch) {
    const int nb = min(a.rows_per_batch, npairs - pb);
    // load real rows (2 per complex line) in digit-reversed z order, with the fused update
    for (int w = threadIdx.x; w < nb * nz; w += blockDim.x) {
      const int lp = w / nz, zz = w - lp * nz;
      const int y0 = 2 * (pb + lp), y1 = y0 + 1;
      T v0, v1 = T(0);
      {
        const int64_t o = fo + (int64_t)y0 * nz + zz;
        v0 = a.r[o];
        if (UPDATE) {
          v0 -= alpha * a.Ap[o];
          a.r_out[o] = v0;
          a.x[o] += alpha * a.p[o];
        }
        acc += (double)v0 * (double)v0;
      }
      if (y1 < ny) {
        const int64_t o = fo + (int64_t)y1 * nz + zz;
        v1 = a.r[o];
        if (UPDATE) {
          v1 -= alpha * a.Ap[o];
          a.r_out[o] = v1;
          a.x[o] += alpha * a.p[o];
        }
        acc += (double)v1 * (double)v1;
      }
      wsp[lp * nz + ipz[zz]] = {v0, v1};
    }
    __syncthreads();
    fft_lds<T, false>(wsp, fz, twz, nb, nz, 1, false, false);
    // unpack the two half spectra into plane rows y0, y1
    for (int w = threadIdx.x; w < nb * nzh; w += blockDim.x) {
      const int lp = w / nzh, k = w - lp * nzh;
      const int y0 = 2 * (pb + lp), y1 = y0 + 1;
      const cplx<T> Z = wsp[lp * nz + k];
      const cplx<T> M = wsp[lp * nz + (k == 0 ? 0 : nz - k)];
      const T h = T(0.5);
      plane[(int64_t)y0 * nzh + k] = {h * (Z.re + M.re), h * (Z.im - M.im)};
      if (y1 < ny) plane[(int64_t)y1 * nzh + k] = {h * (Z.im + M.im), -h * (Z.re - M.re)};
    }
    __syncthreads();
  }
  // y transform: natural in, digit-reversed out
  fft_lds<T, true>(plane, fy, twy, nzh, 1, nzh, false, true);
  // store: position p of the plane holds ky = idx(p)
  cplx<T>* so = a.spec + cd.off_spec + ((int64_t)rhs * 3 + pi.comp) * cd.nspec + (int64_t)pi.x * ny * nzh;
  for (int w = threadIdx.x; w < ny * nzh; w += blockDim.x) {
    const int p = w / nzh, k = w - p * nzh;
    so[(int64_t)py[p] * nzh + k] = plane[w];
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0 && a.part_a) a.part_a[(int64_t)rhs * a.nplanes + item] = t;
}

// Pass B ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) prec_pass_b(const PrecArgs<T> a) {
  extern __shared__ double smem_d[];
  const LineItem li = a.lines[a.line_base + blockIdx.x];
  if (a.done != nullptr && a.done[li.cell]) return;
  const CellDesc cd = a.cells[li.cell];
  const int rhs = blockIdx.y;
  const int nx = cd.nx;
  const int64_t npos = (int64_t)cd.ny * cd.nzh;
  const int len = li.len;
  const int* sl = a.ft.shape_len + 3 * cd.shape_id;
  const FftLen fx = a.ft.lens[sl[0]];
  const cplx<T>* twx = a.ft.tw + fx.tw_off;
  const int* ipx = a.ft.iperm + fx.iperm_off;
  const int* px = a.ft.perm + fx.perm_off;
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_d);  // [nx][3][len]: 3*len lines, stride 1
  cplx<T>* sp = a.spec + cd.off_spec + (int64_t)rhs * 3 * cd.nspec;
  const int nl = 3 * len;

  for (int w = threadIdx.x; w < 3 * nx * len; w += blockDim.x) {
    const int l = w % len;
    const int cx = w / len;
    const int c = cx / nx, xx = cx - c * nx;
    buf[((int64_t)ipx[xx] * 3 + c) * len + l] = sp[(int64_t)c * cd.nspec + (int64_t)xx * npos + li.pos0 + l];
  }
  __syncthreads();
  fft_lds<T, false>(buf, fx, twx, nl, 1, nl, false, true);
  // symbol multiply at natural kx
  const T* sy = a.sym + cd.off_sym;
  for (int w = threadIdx.x; w < nx * len; w += blockDim.x) {
    const int l = w % len, kx = w / len;
    const T* K = sy + ((int64_t)kx * npos + li.pos0 + l) * 6;
    const T k00 = K[0], k11 = K[1], k22 = K[2], k12 = K[3], k02 = K[4], k01 = K[5];
    cplx<T>* b = buf + (int64_t)kx * nl + l;
    const cplx<T> r0 = b[0], r1 = b[len], r2 = b[2 * len];
    b[0] = {k00 * r0.re + k01 * r1.re + k02 * r2.re, k00 * r0.im + k01 * r1.im + k02 * r2.im};
    b[len] = {k01 * r0.re + k11 * r1.re + k12 * r2.re, k01 * r0.im + k11 * r1.im + k12 * r2.im};
    b[2 * len] = {k02 * r0.re + k12 * r1.re + k22 * r2.re, k02 * r0.im + k12 * r1.im + k22 * r2.im};
  }
  __syncthreads();
  fft_lds<T, true>(buf, fx, twx, nl, 1, nl, true, true);
  for (int w = threadIdx.x; w < 3 * nx * len; w += blockDim.x) {
    const int l = w % len;
    const int cp = w / len;
    const int c = cp / nx, p = cp - c * nx;
    sp[(int64_t)c * cd.nspec + (int64_t)px[p] * npos + li.pos0 + l] = buf[((int64_t)p * 3 + c) * len + l];
  }
}

// Pass C ---------------------------------------------------------------------------------------
template <typename T, bool DOT>
__global__ void __launch_bounds__(256) prec_pass_c(const PrecArgs<T> a) {
  extern __shared__ double smem_d[];
  __shared__ double red[4];
  const int item = a.plane_base + blockIdx.x;
  const PlaneItem pi = a.planes[item];
  if (a.done != nullptr && a.done[pi.cell]) return;
  const CellDesc cd = a.cells[pi.cell];
  const int rhs = blockIdx.y;
  const int ny = cd.ny, nz = cd.nz, nzh = cd.nzh;
  const int* sl = a.ft.shape_len + 3 * cd.shape_id;
  const FftLen fy = a.ft.lens[sl[1]];
  const FftLen fz = a.ft.lens[sl[2]];
  const cplx<T>* twy = a.ft.tw + fy.tw_off;
  const cplx<T>* twz = a.ft.tw + fz.tw_off;
  const int* ipz = a.ft.iperm + fz.iperm_off;
  const int* ipy = a.ft.iperm + fy.iperm_off;

  cplx<T>* plane = reinterpret_cast<cplx<T>*>(smem_d);
  cplx<T>* wsp = plane + (int64_t)ny * nzh;
  const cplx<T>* so = a.spec + cd.off_spec + ((int64_t)rhs * 3 + pi.comp) * cd.nspec + (int64_t)pi.x * ny * nzh;
  // load in digit-reversed y order (DIT)
  for (int w = threadIdx.x; w < ny * nzh; w += blockDim.x) {
    const int yy = w / nzh, k = w - yy * nzh;
    plane[(int64_t)ipy[yy] * nzh + k] = so[w];
  }
  __syncthreads();
  fft_lds<T, false>(plane, fy, twy, nzh, 1, nzh, true, true);

  const int64_t nvox = cd.nvox;
  const int64_t fo = cd.off_field + ((int64_t)rhs * 3 + pi.comp) * nvox + (int64_t)pi.x * ny * nz;
  double acc = 0.0;
  const int npairs = (ny + 1) / 2;
  const bool even = (nz % 2) == 0;
  for (int pb = 0; pb < npairs; pb += a.rows_per_batch) {
    const int nb = min(a.rows_per_batch, npairs - pb);
    // Z[k] = A[k] + i B[k] on the full line (Hermitian extension; imaginary parts of the DC and
    // Nyquist terms dropped as irfft does), placed in digit-reversed z order
    for (int w = threadIdx.x; w < nb * nz; w += blockDim.x) {
      const int lp = w / nz, k = w - lp * nz;
      const int y0 = 2 * (pb + lp), y1 = y0 + 1;
      cplx<T> A, B = {T(0), T(0)};
      if (k < nzh) {
        A = plane[(int64_t)y0 * nzh + k];
        if (y1 < ny) B = plane[(int64_t)y1 * nzh + k];
        if (k == 0 || (even && k == nz / 2)) { A.im = T(0); B.im = T(0); }
      } else {
        A = plane[(int64_t)y0 * nzh + (nz - k)];
        A.im = -A.im;
        if (y1 < ny) { B = plane[(int64_t)y1 * nzh + (nz - k)]; B.im = -B.im; }
      }
      wsp[lp * nz + ipz[k]] = {A.re - B.im, A.im + B.re};
    }
    __syncthreads();
    fft_lds<T, false>(wsp, fz, twz, nb, nz, 1, true, false);
    for (int w = threadIdx.x; w < nb * nz; w += blockDim.x) {
      const int lp = w / nz, zz = w - lp * nz;
      const int y0 = 2 * (pb + lp), y1 = y0 + 1;
      const cplx<T> v = wsp[lp * nz + zz];
      {
        const int64_t o = fo + (int64_t)y0 * nz + zz;
        a.z[o] = v.re;
        if (DOT) acc += (double)a.r[o] * (double)v.re;
      }
      if (y1 < ny) {
        const int64_t o = fo + (int64_t)y1 * nz + zz;
        a.z[o] = v.im;
        if (DOT) acc += (double)a.r[o] * (double)v.im;
      }
    }
    __syncthreads();
  }
  if (DOT) {
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) a.part_c[(int64_t)rhs * a.nplanes + item] = t;
  }
}

// Host launchers -------------------------------------------------------------------------------
template <typename T>
size_t prec_plane_lds(int ny, int nz, int rows_per_batch) {
  return ((size_t)ny * (nz / 2 + 1) + (size_t)rows_per_batch * nz) * sizeof(cplx<T>);
}

template <typename T>
hipError_t launch_prec(const PrecArgs<T>& a, int nplanes, int nlines, int update, int dot, size_t lds_plane,
                       size_t lds_line, hipStream_t st) {
  dim3 gp(nplanes, a.m, 1), gl(nlines, a.m, 1);
  if (lds_plane > 65536) {
    (void)hipFuncSetAttribute((const void*)prec_pass_a<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
    (void)hipFuncSetAttribute((const void*)prec_pass_a<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
    (void)hipFuncSetAttribute((const void*)prec_pass_c<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
    (void)hipFuncSetAttribute((const void*)prec_pass_c<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
  }
  if (lds_line > 65536)
    (void)hipFuncSetAttribute((const void*)prec_pass_b<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_line);
  if (update) hipLaunchKernelGGL((prec_pass_a<T, true>), gp, dim3(256), lds_plane, st, a);
  else hipLaunchKernelGGL((prec_pass_a<T, false>), gp, dim3(256), lds_plane, st, a);
  hipLaunchKernelGGL((prec_pass_b<T>), gl, dim3(256), lds_line, st, a);
  if (dot) hipLaunchKernelGGL((prec_pass_c<T, true>), gp, dim3(256), lds_plane, st, a);
  else hipLaunchKernelGGL((prec_pass_c<T, false>), gp, dim3(256), lds_plane, st, a);
  return hipGetLastError();
}

template size_t prec_plane_lds<double>(int, int, int);
template size_t prec_plane_lds<float>(int, int, int);
template hipError_t launch_prec<double>(const PrecArgs<double>&, int, int, int, int, size_t, size_t, hipStream_t);
template hipError_t launch_prec<float>(const PrecArgs<float>&, int, int, int, int, size_t, size_t, hipStream_t);

}  // namespace nlfh
