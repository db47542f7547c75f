/* amr2d_mhd.h -- C ABI of the MI355X-native AMR2D-MHD hot path.
 *
 * The reference (spaceauk/AMR2D_MHD) has no library interface: its hot path is
 * a set of free C++ functions over `meshblock&` called from main.cpp.  Each
 * entry point below names the reference function(s) it replaces (file:line).
 * All pointers are plain host pointers unless stated; the context owns all
 * device memory.  Every call returns an int status (AMR_OK = 0); details of a
 * failure are available from amr_last_error().  Calls on one context must
 * come from one host thread.
 */
#ifndef AMR2D_MHD_H
#define AMR2D_MHD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AMR_OK = 0,
  AMR_ERR_PRESSURE = 1,   /* meshblock.cpp:136-142, 164-170: p <= 0 in the interior */
  AMR_ERR_NAN_VEL = 2,    /* meshblock.cpp:171-176 */
  AMR_ERR_NAN_FLUX = 3,   /* MUSCL2D.cpp:79-85, 103-108 */
  AMR_ERR_CT_FACE = 4,    /* MUSCL2D.cpp:87-91, 110-114 */
  AMR_ERR_NAN_EDGE = 5,   /* celledges.cpp:17-21, 44-47 */
  AMR_ERR_HLLC = 6,       /* riemannS.cpp:269-276 */
  AMR_ERR_DT = 7,         /* main.cpp:235-238 */
  AMR_ERR_OVERFLOW = 8,   /* admesh.cpp:42-45: maxblocks exceeded */
  AMR_ERR_REFINE_NAN = 9, /* refineBi_CT.cpp:47 */
  AMR_ERR_CONFIG = 10,
  AMR_ERR_HIP = 11,
  AMR_ERR_CAPACITY = 12,  /* slot arena of the context exhausted */
  AMR_ERR_HOST = 13,      /* host-side failure (out of host memory): no exception
                             crosses this interface; destroy the context */
  AMR_ERR_IO = 14,        /* checkpoint file missing, unwritable, truncated or corrupt */
  AMR_ERR_CT_PSI = 15     /* CT2D.cpp:60-61: CTtype 2 EMF weight outside [0, 1] */
};

/* slot fields (meshblock arrays, defs.hpp:69-83); layout [slot][var][j][i] */
enum { AMR_U = 0, AMR_US = 1, AMR_BI = 2, AMR_BIS = 3, AMR_FF = 4, AMR_GG = 5, AMR_EMF = 6 };

/* One simulation: the reference's input.txt (nx ny nvar / lenx leny / CFL /
 * IC limiter flux, read at main.cpp:43-69) plus its compile-time constants
 * nghosts and maxlevs (defs.hpp:27-28). */
typedef struct {
  int32_t nx;          /* interior cells per block side: 16 (or 8) */
  int32_t nghosts;     /* 2, or 4 for WENO5 */
  int32_t maxlevs;     /* number of AMR levels, 1..7 */
  int32_t ct_type;     /* corner EMF of CT2D.cpp: 0 or 1 = CTtype 1 (the reference's setting,
                          defs.hpp:34), 2 = CTtype 2, the upwind EMF of CT2D.cpp:28-75 */
  double lenx;         /* domain length parameter (lenx = leny) */
  double cfl;
  char ic[16];         /* SST SST1 SSTy BWx BWx1 BWy BWy1 rotor | rotor_p (parameterised rotor), OT */
  char limiter[16];    /* 1st noSL minmod VanLeer Superbee MC Koren Cada3rd WENO3 WENO5 */
  char flux[8];        /* RUSA HLLE HLLC HLLD */
  double ic_params[8]; /* rotor_p: rho_in, v0, p, Bx, r0, r1, gamma, t_end */
  int32_t periodic;    /* 0: the reference's mirror boundaries (extboundary.cpp); 1: periodic
                          domain [0, lenx]^2 with exact cell size lenx/(2 nx) (build-defined) */
  int32_t rk_stages;   /* 0 or 2: RK2 (timeIntegral.cpp); 3: SSP-RK3 (build-defined); one value
                          per context */
} amr_sim_config;

typedef struct {
  double t, dt, dt_next, gamma, t_end;
  int32_t last_active, nleaf, count, done, status, maxblocks;
} amr_inst_info;

typedef struct {
  int32_t code, inst, nb, i, j, pad;
} amr_diag;

typedef struct amr_ctx amr_ctx;

/* Create a context advancing n_inst independent simulations together (n_inst
 * = 1 is the reference's single run).  slot_capacity bounds the device slot
 * arena over all instances (0: the sum of the instances' maxblocks).  All
 * instances must share nx and nghosts. */
int amr_create(int n_inst, const amr_sim_config *cfgs, int64_t slot_capacity, amr_ctx **out);
int amr_destroy(amr_ctx *ctx);

/* Base grid, initial condition and the initial regrid of every instance:
 * basegrid.cpp:4-89, IC2Dtype.cpp:6-190, main.cpp:97-110 (admesh). */
int amr_init(amr_ctx *ctx);

/* One iteration of the reference time loop (main.cpp:128-247) for every
 * instance that is still running: ghost fills, the two RK2 stages (MUSCL2D,
 * riemannS, RK2ndv1, CT2D), the CFL reduction and the regrid (admesh).
 * *n_active receives the number of instances still running afterwards. */
int amr_step(amr_ctx *ctx, int *n_active);

/* Steps until every instance has reached its end time (main.cpp:128) or
 * max_steps iterations were done (max_steps < 0: no limit).  Returns the
 * number of steps and the leaf-cell updates performed in *zone_updates. */
int amr_run(amr_ctx *ctx, int max_steps, int64_t *steps, int64_t *zone_updates);

/* The two halves of amr_step (no counterpart in the reference): amr_step_begin
 * rebuilds work lists and fill programs after tree changes and launches the
 * step's device work on the context's stream without waiting; amr_step_end
 * waits for it, advances t and dt, regrids and retires finished instances.
 * A context's calls must alternate begin, end. */
int amr_step_begin(amr_ctx *ctx);
int amr_step_end(amr_ctx *ctx, int *n_active);

/* amr_run over several contexts (e.g. an ensemble split into groups), each on
 * its own stream, round robin: while the host completes one context's step
 * and launches its next one, the steps already launched on the others keep
 * the device busy.  Each context runs until its instances are done or fail,
 * or max_steps (< 0: no limit) of its steps; *steps receives the largest
 * step count, *zone_updates the total.  Returns the first error of any
 * context (the others continue unless it is AMR_ERR_HIP).  Results of every
 * instance are those of its own single run. */
int amr_run_group(int n_ctx, amr_ctx **ctxs, int max_steps, int64_t *steps, int64_t *zone_updates);

int amr_info(amr_ctx *ctx, int inst, amr_inst_info *info);
int amr_last_error(amr_ctx *ctx, amr_diag *diag);

/* Copy slots [0, nslots) of one field of one instance to/from host memory,
 * [nb][var][j][i] float64. */
int amr_download(amr_ctx *ctx, int inst, int field, int nslots, double *host);
int amr_upload(amr_ctx *ctx, int inst, int field, int nslots, const double *host);
/* Quadtree of one instance: lp[nb][8], icoord[nb][2], leafs[nb] (defs.hpp:54-58).
 * amr_set_tree rejects (AMR_ERR_CONFIG, instance unchanged) a tree that is not
 * one the reference's bookkeeping can produce (basegrid.cpp:4-89,
 * updatelpup.cpp:4-118, updatelpdown.cpp:5-54): 4 <= nslots <= maxblocks,
 * roots 0..3 at level 0, every other block son lp[nb][3] of its father's group
 * of 4 consecutive ids one level down, consistent sibling pointers inside each
 * group, neighbours in [-1, nslots), refined neighbours of leaves with a son
 * group, leaf flags 0/1, icoord in [0, nx << maxlevs).  amr_validate_tree
 * applies the same check without a context (host only).  amr_download
 * returns zeros for blocks the instance has never used (nslots <= maxblocks);
 * amr_upload accepts only blocks it uses. */
int amr_get_tree(amr_ctx *ctx, int inst, int nslots, int32_t *lp, int32_t *icoord, int32_t *leafs);
int amr_set_tree(amr_ctx *ctx, int inst, int nslots, const int32_t *lp, const int32_t *icoord, const int32_t *leafs);
int amr_validate_tree(int nx, int maxlevs, int nslots, const int32_t *lp, const int32_t *icoord,
                      const int32_t *leafs);
/* t, the dt applied by the current step (meshblock::dt), the dt computed for
 * the next one and the step count: the two-deep dt lag of main.cpp:120-122,
 * 231-234 */
int amr_set_time(amr_ctx *ctx, int inst, double t, double dt, double dt_next, int32_t count);

/* Binary checkpoint / restart of one instance (no counterpart in the
 * reference: its save4py, savedata.cpp:7-46, writes 6-digit ASCII primitives
 * of the leaves and cannot restart a run).  The file holds the configuration,
 * t, dt, dt_next, the step count and status, the quadtree of every allocated
 * block and all seven slot fields of those blocks, with a checksum; a run
 * loaded into a context of the same configuration continues bit for bit as
 * the run that wrote it.  Load fails with AMR_ERR_CONFIG for another
 * configuration and AMR_ERR_IO for a missing, truncated or corrupt file,
 * leaving the instance unchanged. */
int amr_checkpoint_save(amr_ctx *ctx, int inst, const char *path);
int amr_checkpoint_load(amr_ctx *ctx, int inst, const char *path);

/* Individual phases of one step, for phase-level parity checks.
 *  amr_fill:  boundary.cpp:7 (kind 0), boundaryfc.cpp:8 (1), boundaryec.cpp:7 (2)
 *  amr_stage: MUSCL2D.cpp:13 + RK2ndv1 (timeIntegral.cpp:3) over the leaves;
 *             step 1..rk_stages (SSP-RK3 stages 2-3 are build-defined)
 *  amr_ct:    CT2D.cpp:11, same step numbering
 *  amr_cfl:   main.cpp:209-230, *lambda = min over leaves
 *  amr_regrid: admesh.cpp:17 */
int amr_fill(amr_ctx *ctx, int field, int kind);
int amr_stage(amr_ctx *ctx, int step);
int amr_ct(amr_ctx *ctx, int step);
int amr_cfl(amr_ctx *ctx, int inst, double *lambda);
int amr_regrid(amr_ctx *ctx);
/* locate_bounds.cpp:35 for one instance: writes up to cap (source, neighbour,
 * direction) triples, returns the count in *n */
int amr_schedule(amr_ctx *ctx, int inst, int cap, int32_t *ib3, int *n);

/* Pointwise parity entry points on DEVICE pointers (stream may be NULL):
 * riemannS.cpp:7-56 for n face pairs wL[n][8], wR[n][8] -> flux[n][8];
 * solver 0 RUSA 1 HLLE 2 HLLC 3 HLLD, dir 0 x 1 y; status[n] nonzero where
 * the reference throws. */
int amr_riemann_batch(int solver, int dir, double gamma, int n, const double *wL, const double *wR, double *flux,
                      int32_t *status, void *stream);
/* celledges.cpp:5-66 / WENO2D_primvar.cpp:7-198 on nt primitive tiles
 * W[nt][8][NT][NT] -> out[nt][4][8][NT][NT] (wxL wxR wyL wyR), DEVICE pointers */
int amr_recon_batch(int NT, int nghosts, const char *limiter, double dx, int nt, const double *W, double *out,
                    void *stream);

/* Stream of the context (no counterpart in the reference, which runs on the
 * host).  caller = 1: every later operation of the context (steps, phases,
 * uploads, downloads, checkpoints) is enqueued on `stream`, a hipStream_t of
 * the caller (NULL is the null stream), e.g. PyTorch's current stream, so
 * that the caller orders its own work against the context's; caller = 0:
 * back to the context's own stream (the default).  Work enqueued before the
 * switch completes before work enqueued after it.  The caller keeps the
 * stream alive while the context uses it.  Not between amr_step_begin and
 * amr_step_end, nor inside amr_run_group (AMR_ERR_CONFIG). */
int amr_use_stream(amr_ctx *ctx, void *stream, int caller);

/* Graph replay of the step (no counterpart in the reference).  Disabled in
 * this build: amr_set_graph(ctx, 1) returns AMR_ERR_CONFIG; 0 (direct
 * launches, the default) is accepted. */
int amr_set_graph(amr_ctx *ctx, int on);

/* Persistent step kernel (no counterpart in the reference): a step of
 * amr_step / amr_run / amr_run_group runs as ONE cooperative launch over the
 * device work lists of every running instance -- the ghost fills, RK stages,
 * CT kernels and the CFL/criteria reduction of main.cpp:128-230 in order,
 * separated by grid-wide barriers -- instead of one launch per phase.  Same
 * results bit for bit.  mode 1: 256 lanes per workgroup, 2: 512 lanes,
 * 0 and -1 (the default): per-phase launches (graph replay and the
 * stage-kernel tuner apply to these only) -- on MI355X a grid-wide barrier
 * over the 8 XCDs costs more than the kernel boundary it replaces, so per-phase
 * launches are faster for these workloads.  The step kernel runs only where
 * the device keeps its whole grid resident (else per-phase launches); a
 * barrier wait bounded at 2 s that runs out fails the step with AMR_ERR_HIP
 * and switches the context to per-phase launches.  amr_run_group launches per
 * phase, and so do phase-level calls (amr_fill, amr_stage, ...). */
int amr_set_step_kernel(amr_ctx *ctx, int mode);

/* Where the regrid of amr_init / amr_step / amr_run / amr_regrid is decided
 * (admesh.cpp:17-102 with markref.cpp, critneighup.cpp, critneighdown.cpp,
 * updatelpup.cpp and updatelpdown.cpp; the reference runs it on the host).
 * on = 1 (the default): on the device, one workgroup per instance over a
 * device copy of its tree, right after the CFL/criteria kernel of the same
 * step; the host reads back each instance's CFL minimum and the ordered list
 * of its refinements and coarsenings -- no per-leaf values -- and applies
 * that list to its own copy of the tree (which the ghost-fill compiler
 * reads).  on = 0: the host downloads the per-leaf CFL bounds and indicators
 * and decides.  Identical trees and results either way.  Not between
 * amr_step_begin and amr_step_end (AMR_ERR_CONFIG). */
int amr_set_regrid_device(amr_ctx *ctx, int on);

/* Stage-kernel variant (no counterpart in the reference; all variants give
 * identical results): lanes per leaf block and where the face fluxes live
 * for the RK update -- 0: 256 lanes, fluxes in LDS; 1: 512 lanes, LDS;
 * 2: 768 lanes, LDS; 3: 256 lanes, fluxes read back from ff/gg, 4 waves per
 * SIMD; 4: 512 lanes, read back, 4 waves per SIMD; 5: 512 lanes, LDS, 4 waves
 * per SIMD; 6: each leaf split into 2 row strips, one workgroup per strip;
 * 7: 4 row strips (a stage that reads and writes the same field, SSP-RK3's
 * second, runs as variant 2 instead, and the tuner does not try strips with
 * RK3).  Each variant exists generic (solver and limiter read per
 * instance), specialised to HLLD with a MUSCL-family limiter, and, for
 * nghosts 4, specialised to HLLD with WENO5; launches use a specialisation
 * when every running instance has its pair.  -1 (the default)
 * tunes per specialisation and leaf-count bucket (bucket b holds
 * 2^b <= nleaf < 2^(b+1)): the first stage launches in a bucket try the
 * variants in turn, three times each, timed by the kernel's device-clock
 * span, and the one with the smallest median time per leaf block is kept. */
int amr_set_stage_variant(amr_ctx *ctx, int variant);
/* Tuner state of the specialisation in use: chosen[b] (-1 = not decided) and
 * the mean measured microseconds per leaf block of each variant,
 * us_per_leaf[b][6]. */
int amr_stage_tuning(amr_ctx *ctx, int nbuckets, int32_t *chosen, double *us_per_leaf);
/* Import tuner decisions (e.g. from an earlier run of the same workload):
 * chosen[b] >= 0 fixes bucket b's variant for the specialisation the
 * context's running instances select, so later launches in it run no trials;
 * -1 leaves a bucket as it is. */
int amr_set_stage_tuning(amr_ctx *ctx, int nbuckets, const int32_t *chosen);

/* Fill-program cache (no counterpart in the reference): after a tree change
 * an instance reuses a program compiled earlier for the same tree (the key is
 * the whole neighbour table, leaf flags and slot map) instead of compiling
 * and uploading it again.  `entries` programs are kept per instance besides
 * the current one (default 4; 0 disables the cache); changing it drops the
 * cached ones. */
int amr_set_fill_cache(amr_ctx *ctx, int entries);

/* Kernel timing (no counterpart in the reference, which only keeps OpenMP
 * wall-clock accumulators, main.cpp:124-127): launches of the kernel families
 * selected by `mask` (bit 0 stage, 1 ghost fill, 2 CT, 3 CFL/criteria,
 * 4 regrid: the device regrid decision and the refine/coarsen data transfer) are bracketed by HIP events on the context's stream without extra
 * synchronisation (the stage kernel's events are attached to its dispatch);
 * bits 5-7 accumulate host wall time of the topology rebuild (work lists and
 * fill program), of the regrid decisions and of the per-step wait for the
 * device; bit 8 the host time of fill-program compiles, counting compiled
 * programs (cache misses) as launches; bit 9 counts graph replays of the
 * step (no time); bit 11 the host time of graph captures; bit 10 times the persistent
 * step kernel's launches (HIP events around
 * them); with the step kernel, bits 0-3 time its phases from device-clock stamps
 * taken after each phase's closing barrier.  amr_profile_read returns the summed
 * milliseconds and the number of launches (phases, host scopes) of one family
 * since the last amr_profile_enable. */
int amr_profile_enable(amr_ctx *ctx, int mask);
int amr_profile_read(amr_ctx *ctx, int family, double *ms, int64_t *launches);
/* Event sampling of the kernel families (no counterpart in the reference):
 * with stride n > 1 only one launch in n of each family carries events (its
 * time and count go to amr_profile_read), the others run without them; a
 * whole-run mean launch time then costs 1/n of the events' overhead.  1 (the
 * default) times every launch.  Host timers are not sampled. */
int amr_profile_sample(amr_ctx *ctx, int stride);

/* ---- Frozen-mesh step and its discrete adjoint (no counterpart in the
 * reference, which has no adjoint).  One uniform periodic N x N grid of cell
 * size dx, DEVICE pointers: U[8][N][N] (row-major, i fastest), Bx[N][N] (x-face
 * between cells i-1 and i), By[N][N] (y-face between rows j-1 and j).
 * amr_adj_step is one RK2 step of the reference (main.cpp:128-247) on that
 * grid with minmod MUSCL, the `flux` solver ("HLLE" or "HLLD"; NULL = HLLE)
 * and CT, at the given frozen dt; on a single-level mesh it equals the engine
 * bit for bit.
 * amr_adj_step_tape is the same step that also writes the step's tape (the
 * intermediates its reverse needs, amr_adj_tape_doubles(adj) doubles) to
 * `tape` (a DEVICE buffer; NULL: the context's own one).  amr_adj_vjp_tape
 * is the reverse sweep from a tape and the step's input faces: the exact
 * discrete adjoint (U_b, Bx_b, By_b) = J^T (Uo_b, Bxo_b, Byo_b) (branch-wise at
 * limiter and wave-speed switches), without recomputing the step.
 * amr_adj_step_vjp recomputes the step from its input into the context's own
 * tape, then runs the reverse sweep.  Outputs must not alias inputs; calls on
 * one context are stream-ordered and share its work buffers. */
typedef struct amr_adj amr_adj;
int amr_adj_create(int n, double dx, double gamma, const char *flux, amr_adj **out);
int amr_adj_destroy(amr_adj *adj);
int64_t amr_adj_tape_doubles(const amr_adj *adj);
int amr_adj_step(amr_adj *adj, const double *U, const double *Bx, const double *By, double dt, double *Uo,
                 double *Bxo, double *Byo, void *stream);
int amr_adj_step_tape(amr_adj *adj, const double *U, const double *Bx, const double *By, double dt, double *Uo,
                      double *Bxo, double *Byo, double *tape, void *stream);
int amr_adj_vjp_tape(amr_adj *adj, const double *tape, const double *Bx, const double *By, double dt,
                     const double *Uo_b, const double *Bxo_b, const double *Byo_b, double *U_b, double *Bx_b,
                     double *By_b, void *stream);
int amr_adj_step_vjp(amr_adj *adj, const double *U, const double *Bx, const double *By, double dt,
                     const double *Uo_b, const double *Bxo_b, const double *Byo_b, double *U_b, double *Bx_b,
                     double *By_b, void *stream);

/* amr_adj_step without a context and without allocating: no counterpart in
 * the reference.  For callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same step, same results as amr_adj_step on a context created
 * with (n, dx, gamma, flux).  Its work memory is a static device array the
 * library holds on every GPU it runs on until the process ends (25.7 MB), so
 * n is limited to 4..256 (AMR_ERR_CONFIG beyond).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (different streams or
 * host threads) race; calls ordered on one stream do not. */
int amr_adj_step_noalloc(int n, double dx, double gamma, const char *flux, const double *U, const double *Bx,
                         const double *By, double dt, double *Uo, double *Bxo, double *Byo, void *stream);

#ifdef __cplusplus
}
#endif
#endif
