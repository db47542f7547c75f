/* CPU oracle for the AMR2D_MHD per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (spaceauk/AMR2D_MHD), written from the reference's semantics with
 * every floating-point expression evaluated in the reference's order, so that
 * results are bitwise identical to the reference built with -ffp-contract=off.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product (amr2d_mhd_amd) never links
 * or calls it.
 *
 * Layout: every slot field is stored [nb][var][j][i] (i fastest), tiles of
 * NT x NT with NT = nx + 2*nghosts.  Slot ids are the reference's block ids nb.
 */
#ifndef AMR2D_ORACLE_H
#define AMR2D_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_U = 0, ORA_US, ORA_W, ORA_FF, ORA_GG, ORA_BI, ORA_BIS, ORA_EMF, ORA_NFIELD };
enum { ORA_FILL_CC = 0, ORA_FILL_FC = 1, ORA_FILL_EC = 2 };

/* status codes (mirror the reference's throw sites) */
enum {
  ORA_OK = 0,
  ORA_ERR_PRESSURE = 1, /* meshblock.cpp:136-142, 164-170 */
  ORA_ERR_NAN_VEL = 2,  /* meshblock.cpp:171-176 */
  ORA_ERR_NAN_FLUX = 3, /* MUSCL2D.cpp:79-85, 103-108 */
  ORA_ERR_CT_FACE = 4,  /* MUSCL2D.cpp:87-91, 110-114 */
  ORA_ERR_NAN_EDGE = 5, /* celledges.cpp:17-21, 44-47 */
  ORA_ERR_HLLC = 6,     /* riemannS.cpp:269-276 */
  ORA_ERR_DT = 7,       /* main.cpp:235-238 */
  ORA_ERR_OVERFLOW = 8, /* admesh.cpp:42-45 */
  ORA_ERR_REFINE_NAN = 9, /* refineBi_CT.cpp:47 */
  ORA_ERR_CONFIG = 10,
  ORA_ERR_CT_PSI = 15     /* CT2D.cpp:60: upwind EMF weight outside [0, 1] */
};

typedef struct {
  int nx;          /* interior cells per block side (input.txt nx) */
  int nghosts;     /* defs.hpp:27 (2, or 4 for WENO5) */
  int maxlevs;     /* defs.hpp:28 */
  double lenx;     /* input.txt lenx (= leny) */
  double cfl;      /* input.txt CFL */
  char ic[16];     /* IC2Dtype.cpp names: BWx BWx1 BWy BWy1 SST SST1 SSTy rotor; rotor_p, OT */
  char limiter[16];/* slopelimiter.cpp names, WENO3, WENO5 */
  char flux[8];    /* RUSA HLLE HLLC HLLD */
  double ic_params[8]; /* rotor_p (build-defined, no reference counterpart):
                          rho_in, v0, p, Bx, r0, r1, gamma, t_end */
  int periodic;        /* build-defined: wrapped root neighbours, domain [0,lenx]^2 */
  int rk_stages;       /* 0 or 2: timeIntegral.cpp RK2; 3: SSP-RK3 (build-defined) */
  int ct_type;         /* CT2D.cpp corner EMF: 0 or 1 = CTtype 1 (defs.hpp:34, the reference's
                          setting), 2 = CTtype 2 (upwind, CT2D.cpp:28-75) */
} ora_config;

typedef struct ora_dom ora_dom;

ora_dom *ora_create(const ora_config *cfg);
void ora_destroy(ora_dom *d);
int ora_init(ora_dom *d);       /* basegrid + IC + W2U + initial admesh (main.cpp:97-110) */
int ora_step(ora_dom *d);       /* one iteration of main.cpp:128-247 */
int ora_run(ora_dom *d, int max_steps, int32_t *nleaf_log); /* until t>tEnd or max_steps */
int ora_done(const ora_dom *d); /* loop condition main.cpp:128 false */

/* info: lastActive, leaf count, count; t, dom.dt (applied next), dt (lagged next) */
void ora_info(const ora_dom *d, int32_t *ints4, double *dbl4);
uint64_t ora_hash_fields(const ora_dom *d);
uint64_t ora_hash_tree(const ora_dom *d);
int ora_nvar(int field);
int ora_tile(const ora_dom *d);
/* bulk state access, slots [0, n) */
int ora_get_field(const ora_dom *d, int field, int n, double *out);
int ora_set_field(ora_dom *d, int field, int n, const double *in);
int ora_get_tree(const ora_dom *d, int n, int32_t *lp8, int32_t *icoord2, int32_t *leafs);
int ora_set_tree(ora_dom *d, int n, const int32_t *lp8, const int32_t *icoord2, const int32_t *leafs);
void ora_set_time(ora_dom *d, double t, double dom_dt, double dt_next, int count);

/* individual phases (for phase-level parity) */
int ora_locate_bounds(ora_dom *d);
int ora_fill(ora_dom *d, int field, int kind);  /* boundary / boundaryfc / boundaryec */
int ora_stage(ora_dom *d, int step);            /* MUSCL2D + RK over all leaves (step 1..rk) */
int ora_ct(ora_dom *d, int step);               /* CT2D */
int ora_cfl(ora_dom *d, double *lambda);        /* main.cpp:209-230 */
int ora_regrid(ora_dom *d);                     /* admesh */
int ora_nbounds(const ora_dom *d, int32_t *ib3);
void ora_last_error(const ora_dom *d, int32_t *code_slot_i_j);

/* pointwise / block-level functions */
double ora_limiter(const char *name, double r, double eta);
int ora_riemann(const char *method, char dir, double gamma, const double *wL, const double *wR, double *flux);
void ora_wavespeeds(const double *w, double gamma, int norm, double *out4);
/* reconstruction on one W tile -> wxL, wxR, wyL, wyR (each [8][NT][NT]) */
int ora_recon_tile(int NT, int ng, const char *lim, double dx, const double *W, double *wx_l, double *wx_r,
                   double *wy_l, double *wy_r);
/* MUSCL2D on one tile: U (or Us) + Bi (or Bis) -> ff, gg (zero-initialised by caller) */
int ora_muscl_tile(int NT, int ng, const char *lim, const char *flux, double gamma, double dx, int step,
                   const double *U, const double *Bi, double *ff, double *gg);
/* refine (op 0) or coarsen (op 1) on 5 tiles (dad + 4 sons) of U[8] and Bi[2], in place */
int ora_regrid_tiles(int NT, int ng, int op, double gamma, double *U5, double *Bi5);
int ora_criteria_tile(int NT, int ng, int maxlevs, int level, const double *W, int32_t *iref_icoarse);

#ifdef __cplusplus
}
#endif
#endif
