/* Compiled as C11 against include/amr2d_mhd.h and linked with
 * libamr2d_mhd.so by tests/test_library_cpu.py: the header is valid C, every
 * declared entry point resolves at link time, the struct layouts are printed
 * for comparison with the Python binding, and the host-only tree validation
 * runs on a valid and an invalid base tree.  Makes no GPU call. */
#include <stddef.h>
#include <stdio.h>

#include "amr2d_mhd.h"

#define ENTRY(f) (void (*)(void))(f)
#define FIELD(T, m) printf(" %s %zu %zu", #m, offsetof(T, m), sizeof(((T *)0)->m))

int main(void) {
  /* every entry point of the header, by address (a missing export fails the link) */
  void (*const entries[])(void) = {
      ENTRY(amr_create),          ENTRY(amr_destroy),          ENTRY(amr_init),
      ENTRY(amr_step),            ENTRY(amr_run),              ENTRY(amr_step_begin),
      ENTRY(amr_step_end),        ENTRY(amr_run_group),        ENTRY(amr_info),
      ENTRY(amr_last_error),      ENTRY(amr_download),         ENTRY(amr_upload),
      ENTRY(amr_get_tree),        ENTRY(amr_set_tree),         ENTRY(amr_validate_tree),
      ENTRY(amr_set_time),        ENTRY(amr_checkpoint_save),  ENTRY(amr_checkpoint_load),
      ENTRY(amr_fill),            ENTRY(amr_stage),            ENTRY(amr_ct),
      ENTRY(amr_cfl),             ENTRY(amr_regrid),           ENTRY(amr_schedule),
      ENTRY(amr_riemann_batch),   ENTRY(amr_recon_batch),      ENTRY(amr_use_stream),
      ENTRY(amr_set_graph),       ENTRY(amr_set_step_kernel),  ENTRY(amr_set_stage_variant),
      ENTRY(amr_stage_tuning),    ENTRY(amr_set_stage_tuning), ENTRY(amr_set_fill_cache),
      ENTRY(amr_profile_enable),  ENTRY(amr_profile_read),     ENTRY(amr_profile_sample),
      ENTRY(amr_set_regrid_device), ENTRY(amr_adj_create),
      ENTRY(amr_adj_destroy),    ENTRY(amr_adj_tape_doubles), ENTRY(amr_adj_step),
      ENTRY(amr_adj_step_tape),   ENTRY(amr_adj_vjp_tape),     ENTRY(amr_adj_step_vjp),
      ENTRY(amr_adj_step_noalloc),
  };
  size_t n = 0;
  for (size_t k = 0; k < sizeof(entries) / sizeof(entries[0]); k++) n += entries[k] != NULL;
  printf("entries %zu\n", n);

  printf("amr_sim_config %zu", sizeof(amr_sim_config));
  FIELD(amr_sim_config, nx); FIELD(amr_sim_config, nghosts); FIELD(amr_sim_config, maxlevs);
  FIELD(amr_sim_config, ct_type); FIELD(amr_sim_config, lenx); FIELD(amr_sim_config, cfl);
  FIELD(amr_sim_config, ic); FIELD(amr_sim_config, limiter); FIELD(amr_sim_config, flux);
  FIELD(amr_sim_config, ic_params); FIELD(amr_sim_config, periodic); FIELD(amr_sim_config, rk_stages);
  printf("\namr_inst_info %zu", sizeof(amr_inst_info));
  FIELD(amr_inst_info, t); FIELD(amr_inst_info, dt); FIELD(amr_inst_info, dt_next);
  FIELD(amr_inst_info, gamma); FIELD(amr_inst_info, t_end); FIELD(amr_inst_info, last_active);
  FIELD(amr_inst_info, nleaf); FIELD(amr_inst_info, count); FIELD(amr_inst_info, done);
  FIELD(amr_inst_info, status); FIELD(amr_inst_info, maxblocks);
  printf("\namr_diag %zu", sizeof(amr_diag));
  FIELD(amr_diag, code); FIELD(amr_diag, inst); FIELD(amr_diag, nb); FIELD(amr_diag, i);
  FIELD(amr_diag, j); FIELD(amr_diag, pad);
  printf("\n");

  /* 4 root blocks at level 0 (lp[nb] = level, father, first son, son
   * index, 4 neighbours), all leaves, no neighbours */
  int32_t lp[4 * 8], ic[4 * 2], lf[4];
  for (int b = 0; b < 4; b++) {
    for (int s = 0; s < 4; s++) lp[8 * b + s] = 0;
    for (int s = 4; s < 8; s++) lp[8 * b + s] = -1;
    ic[2 * b + 0] = (b % 2) * 16;
    ic[2 * b + 1] = (b / 2) * 16;
    lf[b] = 1;
  }
  printf("validate_base %d\n", amr_validate_tree(16, 4, 4, lp, ic, lf));
  lf[2] = 2; /* a leaf flag other than 0 or 1 */
  printf("validate_bad %d\n", amr_validate_tree(16, 4, 4, lp, ic, lf));
  return 0;
}
