/* nlfh: MI355X (gfx950) small-strain periodic homogenization, C ABI.
 *
 * Drop-in boundary for nlfiber/linear.py (class LinearHomogenization, nlfiber/linear.py:99-264).
 * The reference has no FFI of its own: its interface is that Python class. Each entry point below
 * replaces one of its methods (cited per function); the Python class of this package keeps the
 * reference's signature on top of these calls through a ctypes binding (_lib.py). The finite-strain
 * entry points (nlfh_fs_*) likewise serve the package's counterpart of nlfiber/mech.py.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (HIP), stream-ordered on `stream` (a hipStream_t;
 *    NULL = the null stream) unless marked "host".
 *  - Voxel arrays (rho, s): (nx, ny, nz) C order, per cell, cells concatenated.
 *  - Nodal fields: component-planar (m, 3, nx, ny, nz) per cell, cells concatenated (cell-major).
 *    The reference layout (m, nx, ny, nz, 3) (nlfiber/linear.py:80-92) is this memory viewed
 *    through the axis permutation (0, 2, 3, 4, 1).
 *  - Voigt order xx yy zz yz xz xy, engineering shear; element e = voxel (i,j,k) spans nodes
 *    (i..i+1, j..j+1, k..k+1) mod n (nlfiber/linear.py:21-24).
 *  - Every function returns 0 on success and a nonzero status otherwise; nlfh_last_error()
 *    gives the message (thread-local). No exception crosses the ABI.
 *  - dtype: 0 = float64, 1 = float32.
 */
#ifndef NLFH_H
#define NLFH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nlfh_plan nlfh_plan;

#define NLFH_DTYPE_F64 0
#define NLFH_DTYPE_F32 1

#define NLFH_PREC_FP64 0  /* plain FP64 PCG (the reference's algorithm) */
#define NLFH_PREC_MIXED 1 /* FP32 inner PCG inside FP64 iterative refinement */

/* Message of the last failed call on this thread ("" if none). */
const char* nlfh_last_error(void);

/* ABI version (increases on incompatible change). */
int nlfh_version(void);

/* Plan over `ncells` periodic cells; shapes = ncells x 3 ints (host) (nx, ny, nz).
 * Replaces LinearHomogenization.__init__ (nlfiber/linear.py:102-114): builds the element
 * operator for (E, nu), the preconditioner symbol per distinct shape (nlfiber/linear.py:154-165)
 * and the launch tables. `m_layout` is the number of right-hand sides per cell in batched field
 * arrays (ignored for ncells == 1, where any m <= 6 works). The plan owns its workspaces. */
int nlfh_plan_create(const int* shapes, int ncells, double E, double nu, int m_layout, nlfh_plan** out);
int nlfh_plan_destroy(nlfh_plan* plan);

/* Tiling of the matvec for distinct shape `shape_index` (in order of first appearance):
 * ty node rows per tile, xc x-planes per chunk, zpt z values per lane (1, 2, 3 or 4).
 * Used by the autotuner; defaults come from a heuristic. */
int nlfh_plan_set_tiling(nlfh_plan* plan, int shape_index, int ty, int xc, int zpt);
int nlfh_plan_get_tiling(const nlfh_plan* plan, int shape_index, int* ty, int* xc, int* zpt);

/* Block shapes of the FP32 preconditioner passes of a cubic 16-128 shape: pass_b_positions (ky, kz) positions per
 * block of pass B (the x transforms, Kinv and inverse x transforms), 0 or the default (256 threads; 16 x 16 threads
 * at 128), half or double of it; pass_a_planes x-planes per block of the spectral pass A of the FP32 inner iteration
 * (1, or 2: the second plane's loads in flight during the first's transforms). get returns the values in use and
 * the pass B default (0 for shapes on the generic passes, whose launch shapes are not tunable). No counterpart in
 * the reference (launch-shape choices of LinearHomogenization.precond, nlfiber/linear.py:167-170). */
int nlfh_plan_set_fft_tiling(nlfh_plan* plan, int shape_index, int pass_b_positions, int pass_a_planes);
int nlfh_plan_get_fft_tiling(const nlfh_plan* plan, int shape_index, int* pass_b_positions, int* pass_b_default,
                             int* pass_a_planes);

/* Energy-optimal (Ritz-Galerkin) start from predicted correctors: for every cell and right-hand side
 * I, u_out_I = sum_J a_IJ u_in_J with a_I minimising the PCG energy 1/2 u.Ku + f_I.u over the span
 * of the cell's m predicted fields (G a_I = -(f_I . u_in_J)_J, G_JK = u_in_J . K u_in_K, f = load).
 * A cell whose Gram matrix is singular or non-finite gets the zero start. u_out may equal u_in;
 * coef_out (host, ncells x m x m, nullable) receives a. No counterpart in the reference: it feeds
 * the u0 argument of the solve (nlfiber/linear.py:212-216) from a learned prediction. */
int nlfh_ritz_start(nlfh_plan* plan, const double* s, const double* Es_host, int m, const double* u_in,
                    double* u_out, double* coef_out_host, void* stream);

/* Chebyshev-Jacobi smoothing of a start (no counterpart in the reference: it conditions the learned prediction
 * that feeds the u0 argument of the solve, nlfiber/linear.py:212-216). For every cell and right-hand side I,
 * `degree` steps of the Chebyshev semi-iteration for K u = -load(s, E_I) preconditioned by D = diag K(s), for the
 * spectral interval [lo, hi] of D^-1 K: x_1 = x_0 + r_0 / theta, x_{k+1} = x_k + c1_k (x_k - x_{k-1}) +
 * c2_k D^-1 r_k with r = -(K x + f), in FP32 (one fused operator launch per step). It damps the error components
 * of D^-1 K above lo, which the reference-medium preconditioner resolves slowest in the soft phase. hi <= 0 takes
 * the bound lambda_max(Ke) / Ke_00 of lambda_max(D^-1 K(s)), valid for every s (5.318 at nu = 0.3); hi_used
 * (host, nullable) receives the value used. u_in and u_out are FP64 (m fields) and may be the same array. */
/* lambda_max(Ke) / Ke_00 for material (E, nu): the default hi of nlfh_cheb_smooth (-1 for invalid E, nu). */
double nlfh_jacobi_bound(double E, double nu);
int nlfh_cheb_smooth(nlfh_plan* plan, const double* s, const double* Es_host, int m, const double* u_in, double* u_out,
                     int degree, double lo, double hi, double* hi_used_host, void* stream);

/* Loosest relative tolerance of an FP32 inner solve of NLFH_PREC_MIXED (default 1e-2); each inner
 * solve targets max(inner_tol, 0.5 * tol * ||b|| / ||r||), the reduction its refinement step needs.
 * No counterpart in the reference (which has no mixed precision). */
int nlfh_plan_set_inner_tol(nlfh_plan* plan, double inner_tol);

/* restart = 1: every FP32 inner solve of NLFH_PREC_MIXED restarts its PCG from p = z. restart = 0 (default):
 * plans whose cells all take the FP32 direction iteration (cubic 16-128 cells) carry the last search direction
 * of an inner solve over into the next, after the refinement step has replaced the recursive residual by the
 * true one; other plans always restart. No counterpart in the reference (which has no mixed precision). */
int nlfh_plan_set_inner_restart(nlfh_plan* plan, int restart);

/* s = emin + (1 - emin) clip01(rho)^p over all cells' voxels (s nullable), and optionally
 * dsdrho = p (1 - emin) clip01(rho)^(p-1) inside [0, 1] (bounds included) and 0 outside, the
 * derivative of the reference's clip01 (nlfiber/linear.py:27-30).
 * Replaces LinearHomogenization.stiffness_scale (nlfiber/linear.py:137-141). */
int nlfh_stiffness_scale(nlfh_plan* plan, const double* rho, double emin, double p, double* s, double* dsdrho,
                         void* stream);

/* Ku = K(s) u for m right-hand sides; u, Ku of `dtype`; s is float64.
 * Replaces LinearHomogenization.apply_K (nlfiber/linear.py:143-144), vmapped over m as in
 * nlfiber/linear.py:175. */
int nlfh_apply_K(nlfh_plan* plan, const double* s, const void* u, void* Ku, int m, int dtype, void* stream);

/* f = assembled a(E_I, .) for the macroscopic strains Es (host, m x 6).
 * Replaces LinearHomogenization.load (nlfiber/linear.py:146-149). */
int nlfh_load(nlfh_plan* plan, const double* s, const double* Es_host, int m, double* f, void* stream);

/* r = K(s) u + load(s, Es), not mean-projected (the reference's residual projects;
 * nlfiber/linear.py:151-152; K u and the load are mean-free analytically). */
int nlfh_residual(nlfh_plan* plan, const double* s, const double* Es_host, const double* u, int m, double* r,
                  void* stream);

/* z = irfftn(Kinv . rfftn(r)) per right-hand side and component (zero mean).
 * Replaces LinearHomogenization.precond (nlfiber/linear.py:167-170). */
int nlfh_precond(nlfh_plan* plan, const void* r, void* z, int m, int dtype, void* stream);

/* Correctors: K u_I = -load(s, E_I) on the mean-free space by PCG.
 * Replaces LinearHomogenization.solve / _solve_impl / _cg (nlfiber/linear.py:173-217).
 * u0: nullable initial guess (mean removed as in nlfiber/linear.py:178). precision:
 * NLFH_PREC_FP64 runs the reference's PCG and stopping rule (every right-hand side iterates until
 * all have ||r|| <= tol ||b||, recursive residual; `it` counts iterations as the reference does);
 * NLFH_PREC_MIXED runs FP32 PCG steps inside FP64 iterative refinement and stops on the true FP64
 * residual ||b - K u|| <= tol ||b|| (it = total inner iterations). Outputs (host): it_out[ncells],
 * converged_out[ncells], relres_out[ncells * m] (||r|| / ||b||, true residual in mixed mode,
 * recursive in FP64 mode). Synchronises with the host while it polls for convergence. */
int nlfh_solve(nlfh_plan* plan, const double* s, const double* Es_host, int m, const double* u0, double* u_out,
               double tol, int maxit, int precision, int* it_out, int* converged_out, double* relres_out,
               void* stream);

/* Same as nlfh_solve with an explicit right-hand side b (m fields) and a zero start:
 * replaces LinearHomogenization._solve_rhs (nlfiber/linear.py:208-210), used by the adjoint of
 * the correctors (nlfiber/linear.py:127-131). */
int nlfh_solve_rhs(nlfh_plan* plan, const double* s, const double* b, int m, double* x_out, double tol, int maxit,
                   int precision, int* it_out, int* converged_out, double* relres_out, void* stream);

/* C_IJ = mean_e s_e b_e(E_I + u_I, E_J + u_J): C_out (device, ncells x m x m).
 * dCds_out (nullable, device): per cell (m, m, nx, ny, nz) = b_e,IJ / N = dC_IJ / ds_e.
 * Replaces LinearHomogenization.bilinear (nlfiber/linear.py:220-229) and gives the gradient that
 * the reference obtains by autodiff through it (nlfiber/linear.py:10-13, 259-264). */
int nlfh_bilinear(nlfh_plan* plan, const double* s, const double* Es_host, const double* u, int m, double* C_out,
                  double* dCds_out, void* stream);

/* v -= per-(cell, rhs, component) spatial mean, in place (m fields).
 * Replaces _proj (nlfiber/linear.py:95-96). */
int nlfh_project_mean(nlfh_plan* plan, double* v, int m, void* stream);

/* Host copy of the real symmetric preconditioner symbol of distinct shape `shape_index`:
 * (nx, ny, nz/2+1, 6) as (k00, k11, k22, k12, k02, k01), not divided by N
 * (the real part of LinearHomogenization.Kinv, nlfiber/linear.py:164-165). kinv_out is host. */
int nlfh_symbol(const nlfh_plan* plan, int shape_index, double* kinv_out);

/* Number of bytes of device memory the plan holds. */
long long nlfh_plan_bytes(const nlfh_plan* plan);

/* ---- Learned warm start: periodic 3D conv network kernels (BF16 MFMA, FP32 accumulation) ----
 * No counterpart in the reference's linear path (its only network, nlfiber/prior/unet.py:83-139, is a
 * generative prior with symmetric padding); these kernels train and run the corrector predictor that
 * gives PCG its initial guess u0 (nlfiber/linear.py:212-216). Activations are channels-last
 * [batch][x][y][z][C] (z fastest), bf16 stored as uint16; channel counts are multiples of 16.
 *
 * out[v][co] = bias[co] + sum_{tap, ci} w[tap][co][ci] in[v + d_tap][ci], periodic, where in is the
 * concatenation of source a (ca channels, read at half resolution when a_half) and b (cb channels);
 * w is [27][coutp][ca + cb]; act 1 applies a leaky ReLU (slope 0.1); out_f (f32) and/or out_h
 * (bf16) receive the first `cout` channels. flip negates the tap offsets. */
int nlfh_net_conv(const unsigned short* a, int ca, int a_half, const unsigned short* b, int cb,
                  const unsigned short* w, const float* bias, int n, int batch, int cout, int coutp, int act, int flip,
                  float* out_f, unsigned short* out_h, void* stream);
/* dw[tap][co][ci] = sum_v g[v][co] in[v + d_tap][ci] (in as for nlfh_net_conv); part is scratch of
 * part_len floats for per-chunk partial sums (fixed-order reduction). */
int nlfh_net_wgrad(const unsigned short* g, int cg, const unsigned short* a, int ca, int a_half,
                   const unsigned short* b, int cb, int n, int batch, float* part, long long part_len, float* dw,
                   void* stream);
/* g = bf16(dy * act'(y)) (act 1: leaky ReLU with y its output, 0: identity) */
int nlfh_net_act_grad(const float* dy, const unsigned short* y, unsigned short* g, long long len, int act,
                      void* stream);
/* db[c] = sum_v dy[v][c] act'(y[v][c]); part (nullable) is scratch of part_len floats for per-block partial sums
 * (fixed-order reduction; with null part one block per channel sums alone) */
int nlfh_net_bias_grad(const float* dy, const unsigned short* y, long long nvox, int c, int act, float* part,
                       long long part_len, float* db, void* stream);
/* 2x2x2 average pooling (bf16) */
int nlfh_net_pool(const unsigned short* in, unsigned short* out, int n, int batch, int c, void* stream);
/* coarse (+)= scale * sum of the 8 fine children of channels [c0, c0 + c) of fine (cf channels) */
int nlfh_net_upsum(const float* fine, int cf, int c0, float* coarse, int n, int batch, int c, float scale, int add,
                   void* stream);
/* fine (+)= scale * coarse[v / 2] */
int nlfh_net_updist(const float* coarse, float* fine, int n, int batch, int c, float scale, int add, void* stream);
/* dst[v][0..c) (+)= src[v][c0..c0+c) (src has cs channels) */
int nlfh_net_slice_add(const float* src, int cs, int c0, float* dst, long long nvox, int c, int add, void* stream);
/* mean squared error: grad = 2 (pred - target) / len, per-block loss partials in loss_part[nparts] */
int nlfh_net_mse(const float* pred, const float* target, float* grad, long long len, double* loss_part, int nparts,
                 void* stream);
/* Learned-start input (no counterpart in the reference, whose network is not part of the path): x[v][0] =
   bf16(2 rho[v] - 1) with rho rounded to float32, channels 1..15 zero (16 bf16 channels per voxel, x 16-byte aligned) */
int nlfh_net_input(const double* rho, long long nvox, unsigned short* x, void* stream);
/* Learned-start read-out (no counterpart in the reference, whose network is not part of the path): dst[b][c][v] =
   scale * out[b][v][c] - mean over v, float64, c < nch of the ldc-wide float32 head, channel-major per cell; part is
   scratch of at least batch * 513 * nch doubles */
int nlfh_net_readout(const float* out, int ldc, int nch, long long nvox, int batch, double scale, double* part,
                     long long part_len, double* dst, void* stream);
/* Adam step (bias-corrected, step >= 1, decoupled weight decay wd) on FP32 weights; refreshes the bf16 copy wh */
int nlfh_net_adam(float* w, const float* g, float* m1, float* m2, unsigned short* wh, long long len, float lr,
                  float b1, float b2, float eps, int step, float wd, void* stream);
/* float32 -> bf16 (round to nearest even) */
int nlfh_net_f2bf(const float* in, unsigned short* out, long long len, void* stream);
/* wt[26 - t][ci][co] = w[t][co][ci] (zero for ci >= cin): weights of the data gradient */
int nlfh_net_wflip(const unsigned short* w, int coutp, int cin, unsigned short* wt, int cinp, void* stream);

/* ---- Density map (nlfiber/density.py:76-91, nlfiber/symmetry.py:26-102) ----------------------------
 * phi: symmetric octant (n_o, n_o, n_o) float64; rho, u: full cell (2 n_o)^3 float64, device memory.
 * Replaces DensityMap.__init__ (nlfiber/density.py:79-84): the periodic hat filter of radius r voxels
 * becomes a direct stencil (its support is |offset| <= r). */
typedef struct nlfh_dmap nlfh_dmap;
int nlfh_dmap_create(int n_o, double r, double v, nlfh_dmap** out);
int nlfh_dmap_destroy(nlfh_dmap* dm);
/* u = exact_symmetrize(F mirror(sym_perm(phi)) / std). Replaces DensityMap.smooth (nlfiber/density.py:80-84). */
int nlfh_dmap_smooth(nlfh_dmap* dm, const double* phi, double* u_out, void* stream);
/* rho = sigmoid(beta (u - eta)) with mean(rho) = v; eta_out (host, nullable; synchronises the stream).
 * Replaces DensityMap.rho / eta and threshold / solve_eta (nlfiber/density.py:43-72, 86-91). */
int nlfh_dmap_rho(nlfh_dmap* dm, const double* phi, double beta, double* rho_out, double* eta_out_host, void* stream);
/* phibar_r = (d rho / d phi)^T rbar_r for k in 1..3 rows (rbar: k full cells, phibar: k octants), with the
 * implicit eta derivative of nlfiber/density.py:66-72: the backward pass JAX derives for DensityMap.rho,
 * used by Fiber.value_jac (nlfiber/fiber.py:35-38). Stateless: recomputes the forward pass. */
int nlfh_dmap_rho_vjp(nlfh_dmap* dm, const double* phi, double beta, const double* rbar, int k, double* phibar_out,
                      void* stream);

/* ---- Finite-strain cell operators (nlfiber/mech.py) -------------------------------------------------
 * FP64, device memory. rho (nx, ny, nz); w, v, grad: (nx, ny, nz, 3) C-order, component fastest.
 * mat = {lam, mu, emin, p, wls_b, wls_r0}; flags bit 0: neo-Hookean (else linear), bit 1: void interpolation;
 * nlfh_fs_eval only: bit 2 set = form only the scalars whose bit 8 + q is set (the others are returned as 0).
 * Fbar: host 3x3 row-major; dFbar: device 3x3 row-major (so that a CG loop can form it from its vectors
 * without a host round trip). work: nlfh_fs_work_doubles(nx, ny, nz) doubles of device memory. */
long long nlfh_fs_work_doubles(int nx, int ny, int nz);
/* sg (4, N) = (s_e / (8N), gamma_e, (ds_e/drho_e) / (8N), dgamma_e/drho_e): the stiffness scale and void
 * interpolation factor of every element and their density derivatives (Cell.scale_of / Cell.gamma_of,
 * mech.py:104-112, with clip01's derivative convention), computed once per density. */
int nlfh_fs_prepare(const int* shape, const double* rho, const double* mat, int flags, double* sg, void* stream);
/* mode 0: scalars[0] = W(w, Fbar) = wq sum_e s_e sum_g psi_eff(Fbar + H_g) (Cell.energy, mech.py:141-154);
 * mode 1: also grad = dW/dw (Cell._g, mech.py:172) and scalars[1..9] = Pbar = dW/dFbar (Cell._P, mech.py:173);
 * mode 2: grad = d^2W/dw dz [v; dFbar] and scalars[1..9] = sum of the stress increments, i.e. the product of
 *         the cell Hessian at the state (w, Fbar) with the direction (v, dFbar) (Mixed.hvp_A with the tangent of
 *         Mixed.tangent, mech.py:227-236; the tangent is applied matrix-free). v may be null (zero);
 * mode 3: grad[e] (N values) = d/drho_e of sum_g wq s_e P_eff(F_g; gamma_e) : (dFbar + grad v at g), the density
 *         coupling of Mixed.grad_rho (mech.py:349-363) and the explicit density derivative of Pbar : dFbar;
 *         scalars unused (may be null). */
/* Second and third derivatives of f(s) = <psi_eff(I + sum_I s_I (e_I + grad chi_I))> at s = 0 from the linear
 * correctors chi (6, nx, ny, nz, 3) of the unit Voigt strains: out[0..21) = d2f_IJ for I <= J (row-major over
 * the upper triangle), out[21..77) = d3f_IJK for I <= J <= K (lexicographic). toec (mech.py:454-490) forms C and
 * the third-order constants from these. work: as for nlfh_fs_eval. */
int nlfh_fs_toec(const int* shape, const double* sg, const double* mat, int flags, const double* chi, double* out,
                 double* work, void* stream);
/* CG steps of Mixed._cg (mech.py:266-283) on len unknowns, scalars as device values: alpha = rz / pAp (0 if
 * pAp <= 0), x_out = x + alpha p, r_out = r - alpha Ap, rn2 = |r_out|^2 (fixed-order sum; work: len / 256 + 1
 * doubles); and p_out = z + (rz_new / rz_old) p (rz_old = 0 taken as 1). */
int nlfh_fs_cg_update(int len, double* x_out, double* r_out, const double* x, const double* r, const double* p,
                      const double* Ap, const double* rz, const double* pAp, double* rn2, double* work, void* stream);
int nlfh_fs_cg_direction(int len, double* p_out, const double* z, const double* p, const double* rz_new,
                         const double* rz_old, void* stream);
/* w (nx, ny, nz, 3) -= its per-component mean, in place: the w part of Mixed.proj (mech.py:236-239).
 * work: as for nlfh_fs_eval (at most its first 3 ceil(N / 256) doubles are overwritten). */
int nlfh_fs_project(const int* shape, double* w, double* work, void* stream);
/* The macro unknowns of Mixed (mech.py:183-199): k <= 9 free components of Fbar at the row-major positions idx
 * (a host array). scatter: dst9 (device, 9 values) = the 3x3 direction with src[j] (device) at idx[j] (distinct)
 * and zeros elsewhere, the dFbar of a Hessian product (Mixed.hvp_A, mech.py:227-233); gather: dst[j] = src9[idx[j]]
 * (the free components of a stress sum, device to device). */
int nlfh_fs_macro_scatter(int k, const int* idx, const double* src, double* dst9, void* stream);
int nlfh_fs_macro_gather(int k, const int* idx, const double* src9, double* dst, void* stream);
/* The CG of Mixed._cg (mech.py:266-283) with its dot products folded into the neighbouring passes. part:
 * nlfh_fs_dot_part_doubles() doubles of device memory. nlfh_fs_project_dot projects w like nlfh_fs_project and
 * writes to part[0..256) the partial sums of y . w over the (N, 3) part (w after the projection); the consumer
 * below adds them in a fixed order, plus the k macro terms (the last k entries of both vectors).
 * nlfh_fs_cg_update_parts: pAp from pap_part (the partials of p . Ap) and the macro terms of p and Ap, stored to
 * pap_out; alpha = rz / pAp (0 if pAp <= 0), x_out = x + alpha p, r_out = r - alpha Ap, and rn2_part
 * (ceil(len / 256) doubles) the partials of |r_out|^2. nlfh_fs_cg_direction_parts: rz_new = r . z from rz_part
 * (the partials of r . z) and the macro terms, p_out = z + (rz_new / rz_old) p (rz_old = 0 taken as 1),
 * rn2_out = |r|^2 from rn2_part; rz_new must not be rz_old (every block reads rz_old). */
long long nlfh_fs_dot_part_doubles(void);
int nlfh_fs_project_dot(const int* shape, double* w, const double* y, double* part, void* stream);
int nlfh_fs_cg_update_parts(int len, int k, double* x_out, double* r_out, const double* x, const double* r,
                            const double* p, const double* Ap, const double* rz, const double* pap_part,
                            double* pap_out, double* rn2_part, void* stream);
int nlfh_fs_cg_direction_parts(int len, int k, double* p_out, const double* z, const double* p, const double* r,
                               const double* rz_old, const double* rz_part, double* rz_new, const double* rn2_part,
                               double* rn2_out, void* stream);
int nlfh_fs_eval(const int* shape, const double* sg, const double* mat, int flags, const double* w,
                 const double* Fbar, const double* v, const double* dFbar, int mode, double* grad, double* scalars,
                 double* work, void* stream);
/* nlfh_fs_eval without the work argument, for callers that cannot allocate device memory (code under HIP graph
 * capture, a Newton-CG loop that must not call hipMalloc between steps, harnesses that forbid allocation). No
 * counterpart in the reference. The workspace is a static array of the library, one per GPU, held until the
 * process ends: 24 N + 10 ceil(N / 32) doubles for cells of N <= nlfh_fs_noalloc_max_voxels() = 32768 voxels
 * (32^3; 6.1 MiB). Larger cells are refused (NLFH_EUNSUPPORTED). The function itself calls no allocator; its
 * first call on a device looks up the array's address there. Every call on one GPU shares the array: calls that
 * overlap in time on one GPU (other streams, other host threads) race; calls ordered on one stream do not.
 * Every value the kernels read from it is written by the same call first, so nothing needs clearing.
 * Results are bitwise those of nlfh_fs_eval (same kernels, same order of sums). */
long long nlfh_fs_noalloc_max_voxels(void);
int nlfh_fs_eval_noalloc(const int* shape, const double* sg, const double* mat, int flags, const double* w,
                         const double* Fbar, const double* v, const double* dFbar, int mode, double* grad,
                         double* scalars, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NLFH_H */
