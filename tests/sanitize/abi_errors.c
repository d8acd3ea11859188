/* Host-side checks of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer: every call below
 * either does host arithmetic only or must be refused by the library's argument checks before any device
 * work, so the program runs without a GPU. Exit status 0 when every refusal and value is as expected. */
#include <stdio.h>
#include <string.h>

#include "nlfh.h"

static int fails = 0;
#define EXPECT(cond, what)                                     \
  do {                                                         \
    if (!(cond)) {                                             \
      fprintf(stderr, "FAILED: %s (%s)\n", what, nlfh_last_error()); \
      ++fails;                                                 \
    }                                                          \
  } while (0)

int main(void) {
  EXPECT(nlfh_version() == 1, "version");
  nlfh_plan* plan = (nlfh_plan*)0;
  int shapes[6] = {8, 8, 8, 7, 8, 8};
  EXPECT(nlfh_plan_create(NULL, 1, 1.0, 0.3, 6, &plan) != 0 && strstr(nlfh_last_error(), "ncells"), "null shapes");
  EXPECT(nlfh_plan_create(shapes, 0, 1.0, 0.3, 6, &plan) != 0, "zero cells");
  EXPECT(nlfh_plan_create(shapes, 1, 1.0, 0.6, 6, &plan) != 0 && strstr(nlfh_last_error(), "nu"), "nu >= 0.5");
  EXPECT(nlfh_plan_create(shapes, 1, 1.0, 0.3, 7, &plan) != 0 && strstr(nlfh_last_error(), "m_layout"), "m_layout");
  EXPECT(nlfh_plan_create(shapes, 1, 1.0, 0.3, 6, NULL) != 0, "null out");
  EXPECT(plan == NULL, "no plan on failure");
  EXPECT(nlfh_plan_bytes(NULL) == 0, "bytes of null plan");
  double Es[6] = {1, 0, 0, 0, 0, 0};
  int it = 0, cv = 0;
  double rr = 0.0;
  EXPECT(nlfh_solve(NULL, NULL, Es, 1, NULL, NULL, 1e-8, 10, 0, &it, &cv, &rr, NULL) != 0, "solve without plan");
  EXPECT(nlfh_cheb_smooth(NULL, NULL, Es, 1, NULL, NULL, 8, 0.1, 1.0, NULL, NULL) != 0, "smooth without plan");
  EXPECT(nlfh_ritz_start(NULL, NULL, Es, 1, NULL, NULL, NULL, NULL) != 0, "ritz without plan");
  EXPECT(nlfh_plan_set_inner_restart(NULL, 0) != 0, "inner restart without plan");
  EXPECT(nlfh_plan_set_fft_tiling(NULL, 0, 32, 1) != 0, "fft tiling without plan");
  /* host arithmetic: the spectral bound of the Jacobi-scaled operator (24x24 element matrix, power iteration) */
  const double b = nlfh_jacobi_bound(1.0, 0.3);
  EXPECT(b > 5.31 && b < 5.33, "jacobi bound at nu = 0.3");
  EXPECT(nlfh_jacobi_bound(1.0, 0.5) == -1.0, "jacobi bound rejects nu = 0.5");
  /* finite-strain sizes and refusals */
  EXPECT(nlfh_fs_work_doubles(64, 64, 64) == 24LL * 262144 + (262144 / 32) * 10, "fs work size");
  EXPECT(nlfh_fs_work_doubles(0, 4, 4) < 0, "fs work size of an empty cell");
  EXPECT(nlfh_fs_noalloc_max_voxels() == 32768, "no-alloc limit");
  int sh[3] = {4, 4, 4};
  double mat[6] = {0.5769, 0.3846, 1e-3, 3.0, 500.0, 0.01};
  double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  EXPECT(nlfh_fs_eval(sh, NULL, mat, 1, NULL, F, NULL, NULL, 1, NULL, NULL, NULL, NULL) != 0, "fs_eval nulls");
  EXPECT(nlfh_fs_eval(sh, NULL, mat, 1, NULL, F, NULL, NULL, 7, NULL, NULL, NULL, NULL) != 0, "fs_eval mode");
  EXPECT(nlfh_fs_prepare(NULL, NULL, mat, 1, NULL, NULL) != 0, "fs_prepare nulls");
  EXPECT(nlfh_fs_project(sh, NULL, NULL, NULL) != 0, "fs_project nulls");
  int midx[3] = {4, 9, 4};
  double d9[9];
  EXPECT(nlfh_fs_macro_scatter(10, midx, d9, d9, NULL) != 0, "macro scatter k > 9");
  EXPECT(nlfh_fs_macro_scatter(2, midx, d9, d9, NULL) != 0, "macro scatter index out of range");
  midx[1] = 0;
  EXPECT(nlfh_fs_macro_scatter(3, midx, d9, d9, NULL) != 0, "macro scatter repeated index");
  EXPECT(nlfh_fs_macro_gather(1, NULL, d9, d9, NULL) != 0, "macro gather without idx");
  EXPECT(nlfh_fs_macro_gather(-1, midx, d9, d9, NULL) != 0, "macro gather k < 0");
  EXPECT(nlfh_fs_dot_part_doubles() == 256 + 3 * 512, "dot partials size");
  EXPECT(nlfh_fs_project_dot(sh, d9, NULL, d9, NULL) != 0, "project_dot without y");
  EXPECT(nlfh_fs_cg_update_parts(8, 9, d9, d9, d9, d9, d9, d9, d9, d9, d9, d9, NULL) != 0, "cg update parts k > len");
  EXPECT(nlfh_fs_cg_direction_parts(8, 0, d9, d9, d9, d9, d9, d9, d9, d9, d9, NULL) != 0, "cg direction rz_new == rz_old");
  EXPECT(nlfh_fs_cg_update(0, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL) != 0, "cg update");
  /* network kernels: channel counts not multiples of 16 */
  EXPECT(nlfh_net_conv(NULL, 15, 0, NULL, 0, NULL, NULL, 8, 1, 16, 16, 1, 0, NULL, NULL, NULL) != 0, "conv args");
  EXPECT(nlfh_net_wgrad(NULL, 16, NULL, 15, 0, NULL, 0, 8, 1, NULL, 0, NULL, NULL) != 0, "wgrad args");
  EXPECT(nlfh_net_bias_grad(NULL, NULL, 1, 16, 0, NULL, 0, NULL, NULL) != 0, "bias grad args");
  EXPECT(nlfh_net_readout(NULL, 32, 18, 8, 1, 4.0, NULL, 0, NULL, NULL) != 0, "readout nulls");
  EXPECT(nlfh_net_readout((const float*)d9, 16, 18, 8, 1, 4.0, d9, 1, d9, NULL) != 0 &&
             strstr(nlfh_last_error(), "ldc >= nch"), "readout ldc < nch");
  EXPECT(nlfh_net_input(NULL, 8, NULL, NULL) != 0, "input nulls");
  if (fails == 0) printf("abi_errors: all checks passed\n");
  return fails == 0 ? 0 : 1;
}
