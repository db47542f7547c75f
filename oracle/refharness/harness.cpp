// Test-only driver for the compiled reference (AMR2D_MHD).  Built by
// oracle/build_ref.sh from the reference sources where they lie; the
// outputs go to oracle/_ref/ and are used ONLY as the parity checker and as
// the CPU-baseline timer.  Nothing in the product links this.
//
// Design: the reference's own main() (main.cpp) runs unmodified.  Its phase
// functions are renamed at compile time (-Dboundary=ref_boundary, ...) so
// that the calls main.cpp makes land in the thin wrappers below, which record
// state hashes / dumps and then call the renamed reference function.  This
// keeps the reference's time loop (main.cpp:128-247) byte-for-byte its own.
//
// Modes
//   run  : run the reference simulation (input.txt written into --dir),
//          log per-step hashes, dump states, optional per-phase trace.
//   vec  : evaluate single reference functions on vectors read from a file.
#include "defs.hpp"
#include "timeIntegral.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

// ---- reference entry points (renamed definitions; see build_ref.sh) -------
int ref_main();
void ref_boundary(meshblock &dom, real ****Q, int nvar);
void ref_boundaryfc(meshblock &dom, real ****Bi, int nvar);
void ref_boundaryec(meshblock &dom, real ****E, int nvar);
void ref_CT2D(meshblock &dom, real ****q, real ****Bi, int step, real ****Binew);
void ref_MUSCL2D(meshblock &dom, int nb, int step);
void ref_admesh(meshblock &dom);
void ref_locate_bounds(meshblock &dom);
void ref_save4py(meshblock &dom, int id, real t);
void ref_criteria(meshblock &dom, int nb);
// reference functions called directly in vec mode
real slopelimiter(string limiter, real r, real eta);
void riemannS(string fluxMth, vector<real> wL, vector<real> wR, real gamma, char direc, vector<real> &flux);
void wavespeeds(vector<real> w, real gamma, int norm, real &a, real &ca, real &can, real &cf);
void celledges(meshblock &dom, int nb, int step);
void WENO2D_3rd(meshblock &dom, string direc, int nb);
void WENO2D_5th(meshblock &dom, string direc, int nb);
void refine(meshblock &dom, real ****Q, int nvar, int nb, int s1, int s2, int s3, int s4);
void coarsen(meshblock &dom, real ****Q, int nvar, int dad, int s1, int s2, int s3, int s4);
void refineBi_CT(meshblock &dom, real ****Bi, int nvar, int nb, int s1, int s2, int s3, int s4);
void coarsenBi_CT(meshblock &dom, real ****Bi, int nvar, int dad, int s1, int s2, int s3, int s4);

// main.cpp calls system() to launch its plotting script; the link uses
// --wrap=system so that call lands here and is ignored.
extern "C" int __wrap_system(const char *) { return 0; }

namespace {

struct Opts {
  std::string dir = ".";
  int steps = -1;          // stop after this many completed steps (-1: run to tEnd)
  std::string logf;        // per-step binary log
  std::vector<int> dumps;  // dump full state at the end of these steps
  int trace = -1;          // dump state around every phase of this step
  std::string timingf;     // wall-clock of the time loop
  std::string steptimesf;  // per step: seconds since the loop started at its end, leaves it advanced
  // random-tree mode (criteria replaced by seeded random marks)
  unsigned long long seed = 0;
  double p_ref = 0.0, p_coarse = 0.0;
  int rand_rounds = 0;
} g;

meshblock *D = nullptr;
int g_step = -1;           // index of the step in progress (0-based)
int g_admesh_calls = 0;
int g_event = 0;
double g_t = 0.0;          // t reconstructed from the applied dt sequence
bool g_zeroed = false;
int g_muscl_stage_seen[3] = {-1, -1, -1};
double g_dt_applied = 0.0;
FILE *g_log = nullptr;
FILE *g_steptimes = nullptr;
std::chrono::steady_clock::time_point g_t0;
bool g_t0_set = false;

struct StopRun {};

const uint64_t FNV_OFF = 0xcbf29ce484222325ULL, FNV_PRIME = 0x100000001b3ULL;
inline void fnv(uint64_t &h, uint64_t w) { h ^= w; h *= FNV_PRIME; }
inline uint64_t bits(double x) { uint64_t u; std::memcpy(&u, &x, 8); return u; }

// field order used by hashes and dumps: U Us Bi Bis ff gg EMF (W dumped separately)
void field_list(std::vector<std::pair<real ****, int>> &f) {
  f = {{D->U, 8}, {D->Us, 8}, {(real ****)D->Bi, 2}, {(real ****)D->Bis, 2},
       {D->ff, 8}, {D->gg, 8}, {(real ****)D->EMF, 1}};
}

uint64_t hash_fields() {
  std::vector<std::pair<real ****, int>> f; field_list(f);
  uint64_t h = FNV_OFF;
  for (auto &p : f)
    for (int nb = 0; nb < D->lastActive; nb++)
      for (int k = 0; k < p.second; k++)
        for (int j = 0; j < D->ny; j++)
          for (int i = 0; i < D->nx; i++) fnv(h, bits(p.first[i][j][k][nb]));
  return h;
}

int count_leaves() {
  int n = 0;
  for (int nb = 0; nb < D->lastActive; nb++) n += D->leafs[nb] ? 1 : 0;
  return n;
}

uint64_t hash_tree() {
  uint64_t h = FNV_OFF;
  for (int nb = 0; nb < D->lastActive; nb++) {
    for (int q = 0; q < 8; q++) fnv(h, (uint64_t)(int64_t)D->lp[nb][q]);
    fnv(h, (uint64_t)(int64_t)D->icoord[nb][0]);
    fnv(h, (uint64_t)(int64_t)D->icoord[nb][1]);
    fnv(h, (uint64_t)(D->leafs[nb] ? 1 : 0));
  }
  return h;
}

// State file: int32 header[8] = {magic, version, NT, nghosts, lastActive,
// nbleafs, step, maxlevs}; double t, dt; int32 lp[nA][8], icoord[nA][2],
// leafs[nA]; then float64 fields U Us W Bi Bis ff gg EMF as [nb][k][j][i].
void dump_state(const std::string &path, double t, double dt) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) { std::perror(path.c_str()); std::exit(3); }
  int nA = D->lastActive;
  int32_t hdr[8] = {0x53524d41, 1, D->nx, nghosts, nA, D->nbleafs, g_step, maxlevs};
  std::fwrite(hdr, 4, 8, fp);
  double td[2] = {t, dt};
  std::fwrite(td, 8, 2, fp);
  std::vector<int32_t> iv;
  for (int nb = 0; nb < nA; nb++) for (int q = 0; q < 8; q++) iv.push_back(D->lp[nb][q]);
  for (int nb = 0; nb < nA; nb++) { iv.push_back(D->icoord[nb][0]); iv.push_back(D->icoord[nb][1]); }
  for (int nb = 0; nb < nA; nb++) iv.push_back(D->leafs[nb] ? 1 : 0);
  std::fwrite(iv.data(), 4, iv.size(), fp);
  std::vector<std::pair<real ****, int>> f = {
      {D->U, 8}, {D->Us, 8}, {D->W, 8}, {(real ****)D->Bi, 2}, {(real ****)D->Bis, 2},
      {D->ff, 8}, {D->gg, 8}, {(real ****)D->EMF, 1}};
  std::vector<double> buf((size_t)D->nx * D->ny);
  for (auto &p : f)
    for (int nb = 0; nb < nA; nb++)
      for (int k = 0; k < p.second; k++) {
        for (int j = 0; j < D->ny; j++)
          for (int i = 0; i < D->nx; i++) buf[(size_t)j * D->nx + i] = p.first[i][j][k][nb];
        std::fwrite(buf.data(), 8, buf.size(), fp);
      }
  std::fclose(fp);
}

void trace(const char *what, const char *arr, const char *when) {
  if (g_step != g.trace) return;
  char name[256];
  std::snprintf(name, sizeof name, "%s/trace_%03d_%s_%s_%s.bin", g.dir.c_str(), g_event++, what, arr, when);
  dump_state(name, g_t, D->dt);
}

const char *arrname(void *Q) {
  if (Q == (void *)D->U) return "U";
  if (Q == (void *)D->Us) return "Us";
  if (Q == (void *)D->Bi) return "Bi";
  if (Q == (void *)D->Bis) return "Bis";
  if (Q == (void *)D->ff) return "ff";
  if (Q == (void *)D->gg) return "gg";
  if (Q == (void *)D->EMF) return "EMF";
  return "X";
}

// Make the reference's implicit zero-initialisation explicit: every slot array
// and tree entry the reference has not written yet is set to zero/false, and
// the (otherwise uninitialised) applied dt of step 0 is 0.
void zero_unwritten() {
  int nA = D->lastActive;  // roots 0..nA-1 hold the IC in U, W and Bi
  int nx = D->nx, ny = D->ny;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      for (int k = 0; k < D->nvar; k++)
        for (int nb = 0; nb < maxblocks; nb++) {
          if (nb >= nA) { D->U[i][j][k][nb] = 0; D->W[i][j][k][nb] = 0; }
          D->Us[i][j][k][nb] = 0; D->ff[i][j][k][nb] = 0; D->gg[i][j][k][nb] = 0;
        }
      for (int k = 0; k < 2; k++)
        for (int nb = 0; nb < maxblocks; nb++) {
          if (nb >= nA) D->Bi[i][j][k][nb] = 0;
          D->Bis[i][j][k][nb] = 0;
        }
      for (int nb = 0; nb < maxblocks; nb++) D->EMF[i][j][0][nb] = 0;
    }
  for (int nb = nA; nb < maxblocks; nb++) {
    for (int q = 0; q < 8; q++) D->lp[nb][q] = 0;
    D->icoord[nb][0] = D->icoord[nb][1] = 0;
    D->leafs[nb] = false; D->iref[nb] = false; D->icoarse[nb] = false;
  }
  for (int nb = 0; nb < nA; nb++) D->icoarse[nb] = false;
  D->dt = 0;
}

// One record per completed step (step=-1: state after the initial regrid).
// nleaf_* count leafs[] directly (the reference's nbleafs is a sizing bound,
// updatelpup.cpp:109 adds 4 per refinement).
struct LogRec {
  int32_t step, lastActive_pre, nleaf_pre, lastActive_post;
  int32_t nleaf_post, pad;
  double dt_applied, t;
  uint64_t hf_pre, ht_pre, hf_post, ht_post;
};

// Random-tree mode: grow/shrink the tree with seeded random marks through the
// reference's own admesh, then overwrite every slot field with seeded values
// and run the reference ghost fills and one CT pass on that state.
double rnd() {
  g.seed = g.seed * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(g.seed >> 11) * (1.0 / 9007199254740992.0);
}

void random_tree_fills(meshblock &dom) {
  for (int r = 0; r < g.rand_rounds; r++) ref_admesh(dom);
  std::vector<std::pair<real ****, int>> f; field_list(f);
  for (auto &p : f)
    for (int nb = 0; nb < dom.lastActive; nb++)
      for (int k = 0; k < p.second; k++)
        for (int j = 0; j < dom.ny; j++)
          for (int i = 0; i < dom.nx; i++) p.first[i][j][k][nb] = 0.5 + rnd();
  dom.dt = 1e-3;
  g_step = 0;
  dump_state("state_rt_pre.bin", 0.0, dom.dt);
  ref_locate_bounds(dom);
  ref_boundary(dom, dom.U, 8);
  ref_boundaryfc(dom, (real ****)dom.Bi, 2);
  ref_boundaryfc(dom, dom.ff, 8);
  ref_boundaryfc(dom, dom.gg, 8);
  ref_boundaryec(dom, (real ****)dom.EMF, 1);
  dump_state("state_rt_fill.bin", 0.0, dom.dt);
  ref_CT2D(dom, dom.Us, (real ****)dom.Bi, 1, (real ****)dom.Bis);
  dump_state("state_rt_ct.bin", 0.0, dom.dt);
}

}  // namespace

// ---- wrappers the reference main() calls -----------------------------------
void locate_bounds(meshblock &dom) {
  D = &dom;
  g_step++;
  g_event = 0;
  g_muscl_stage_seen[1] = g_muscl_stage_seen[2] = -1;
  if (!g_t0_set) { g_t0 = std::chrono::steady_clock::now(); g_t0_set = true; }
  g_dt_applied = dom.dt;
  ref_locate_bounds(dom);
}

void boundary(meshblock &dom, real ****Q, int nvar) {
  D = &dom;
  trace("fillcc", arrname(Q), "pre");
  ref_boundary(dom, Q, nvar);
  trace("fillcc", arrname(Q), "post");
}

void boundaryfc(meshblock &dom, real ****Q, int nvar) {
  D = &dom;
  trace("fillfc", arrname(Q), "pre");
  ref_boundaryfc(dom, Q, nvar);
  trace("fillfc", arrname(Q), "post");
}

void boundaryec(meshblock &dom, real ****Q, int nvar) {
  D = &dom;
  trace("fillec", arrname(Q), "pre");
  ref_boundaryec(dom, Q, nvar);
  trace("fillec", arrname(Q), "post");
}

void CT2D(meshblock &dom, real ****q, real ****Bi, int step, real ****Binew) {
  D = &dom;
  trace("ct", step == 1 ? "s1" : "s2", "pre");
  ref_CT2D(dom, q, Bi, step, Binew);
  trace("ct", step == 1 ? "s1" : "s2", "post");
}

void MUSCL2D(meshblock &dom, int nb, int step) {
  D = &dom;
  if (g_muscl_stage_seen[step] != g_step) {
    g_muscl_stage_seen[step] = g_step;
    trace("stage", step == 1 ? "s1" : "s2", "pre");
  }
  ref_MUSCL2D(dom, nb, step);
}

void save4py(meshblock &, int, real) {}

void criteria(meshblock &dom, int nb) {
  if (g.rand_rounds <= 0) { ref_criteria(dom, nb); return; }
  // random marks reproducing criteria's outcome semantics (criteria.cpp:151-163)
  g.seed = g.seed * 6364136223846793005ULL + 1442695040888963407ULL;
  double u = (double)(g.seed >> 11) * (1.0 / 9007199254740992.0);
  if (u < g.p_ref) {
    if (dom.lp[nb][0] < maxlevs - 1) { dom.iref[nb] = true; dom.icoarse[nb] = false; }
  } else if (u < g.p_ref + g.p_coarse) {
    dom.icoarse[nb] = true;
  }
}

void admesh(meshblock &dom) {
  D = &dom;
  if (!g_zeroed) {  // initial regrid at main.cpp:110
    zero_unwritten();
    g_zeroed = true;
    if (g.rand_rounds > 0) { random_tree_fills(dom); throw StopRun(); }
    ref_admesh(dom);
    if (g_log) {
      LogRec r{};
      r.step = -1; r.lastActive_post = dom.lastActive; r.nleaf_post = count_leaves();
      r.hf_post = hash_fields(); r.ht_post = hash_tree();
      std::fwrite(&r, sizeof r, 1, g_log);
    }
    for (int s : g.dumps)
      if (s == -1) dump_state(g.dir + "/state_init.bin", 0.0, 0.0);
    return;
  }
  // end of step g_step: dom.dt is the dt applied in this step
  g_t = g_t + dom.dt;
  LogRec r{};
  r.step = g_step; r.dt_applied = g_dt_applied; r.t = g_t;
  r.lastActive_pre = dom.lastActive; r.nleaf_pre = count_leaves();
  if (g_log) { r.hf_pre = hash_fields(); r.ht_pre = hash_tree(); }
  trace("regrid", "all", "pre");
  ref_admesh(dom);
  trace("regrid", "all", "post");
  r.lastActive_post = dom.lastActive; r.nleaf_post = count_leaves();
  if (g_steptimes) {
    // end of the loop iteration (main.cpp:128-247): the regrid was its last work
    const double rec[2] = {std::chrono::duration<double>(std::chrono::steady_clock::now() - g_t0).count(),
                           (double)r.nleaf_pre};
    std::fwrite(rec, sizeof rec, 1, g_steptimes);
  }
  if (g_log) { r.hf_post = hash_fields(); r.ht_post = hash_tree(); std::fwrite(&r, sizeof r, 1, g_log); std::fflush(g_log); }
  for (int s : g.dumps)
    if (s == g_step) {
      char name[64]; std::snprintf(name, sizeof name, "/state_%05d.bin", s);
      dump_state(g.dir + name, g_t, dom.dt);
    }
  if (g.steps >= 0 && g_step + 1 >= g.steps) throw StopRun();
}

// ---- vec mode ----------------------------------------------------------------
namespace {

std::vector<char> slurp(const char *path) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) { std::perror(path); std::exit(3); }
  std::vector<char> b;
  char tmp[1 << 16];
  size_t n;
  while ((n = std::fread(tmp, 1, sizeof tmp, fp)) > 0) b.insert(b.end(), tmp, tmp + n);
  std::fclose(fp);
  return b;
}

struct Reader {
  const std::vector<char> &b; size_t off = 0;
  template <class T> T get() { T v; std::memcpy(&v, b.data() + off, sizeof(T)); off += sizeof(T); return v; }
  std::string str(int n) { std::string s(b.data() + off, b.data() + off + n); off += n; s = s.c_str(); return s; }
  const double *dptr(size_t n) { const double *p = (const double *)(b.data() + off); off += 8 * n; return p; }
};

void put(FILE *fp, const void *p, size_t n) { std::fwrite(p, 1, n, fp); }

// Block-level setup: one meshblock with NT x NT tiles, slots for the blocks.
void block_setup(meshblock &dom, int NT, double gamma, const std::string &lim, const std::string &flux) {
  dom.setSize(NT, NT, 8, 1.0, 1.0);
  dom.setType("BWx", lim, flux);
  dom.setParam(gamma, 1.0);
  for (int nb = 0; nb < maxblocks; nb++) {
    for (int q = 0; q < 8; q++) dom.lp[nb][q] = 0;
    dom.icoord[nb][0] = dom.icoord[nb][1] = 0;
    dom.leafs[nb] = true; dom.iref[nb] = false; dom.icoarse[nb] = false;
  }
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++)
      for (int k = 0; k < 8; k++) {
        dom.wxL[i][j][k] = dom.wxR[i][j][k] = dom.wyL[i][j][k] = dom.wyR[i][j][k] = 0;
        dom.dwdx[i][j][k] = dom.dwdy[i][j][k] = 0;
        for (int nb = 0; nb < maxblocks; nb++) {
          dom.U[i][j][k][nb] = dom.Us[i][j][k][nb] = dom.W[i][j][k][nb] = 0;
          dom.ff[i][j][k][nb] = dom.gg[i][j][k][nb] = 0;
          if (k < 2) { dom.Bi[i][j][k][nb] = 0; dom.Bis[i][j][k][nb] = 0; }
          if (k < 1) dom.EMF[i][j][k][nb] = 0;
        }
      }
}

void tile_in(real ****A, int nvar, int nb, int NT, const double *src) {
  for (int k = 0; k < nvar; k++)
    for (int j = 0; j < NT; j++)
      for (int i = 0; i < NT; i++) A[i][j][k][nb] = src[((size_t)k * NT + j) * NT + i];
}
void tile_out(real ****A, int nvar, int nb, int NT, FILE *fp) {
  std::vector<double> b((size_t)nvar * NT * NT);
  for (int k = 0; k < nvar; k++)
    for (int j = 0; j < NT; j++)
      for (int i = 0; i < NT; i++) b[((size_t)k * NT + j) * NT + i] = A[i][j][k][nb];
  put(fp, b.data(), 8 * b.size());
}
void scratch_out(real ***A, int NT, FILE *fp) {
  std::vector<double> b((size_t)8 * NT * NT);
  for (int k = 0; k < 8; k++)
    for (int j = 0; j < NT; j++)
      for (int i = 0; i < NT; i++) b[((size_t)k * NT + j) * NT + i] = A[i][j][k];
  put(fp, b.data(), 8 * b.size());
}
void scratch_zero(meshblock &dom, int NT) {
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++)
      for (int k = 0; k < 8; k++) dom.wxL[i][j][k] = dom.wxR[i][j][k] = dom.wyL[i][j][k] = dom.wyR[i][j][k] = 0;
}

int vec_main(const char *kind, const char *inp, const char *outp) {
  std::vector<char> buf = slurp(inp);
  Reader rd{buf};
  FILE *fo = std::fopen(outp, "wb");
  if (!fo) { std::perror(outp); return 3; }
  std::string k = kind;
  if (k == "limiter") {
    int n = rd.get<int32_t>();
    std::string name = rd.str(32);
    const double *r = rd.dptr(n), *eta = rd.dptr(n);
    for (int q = 0; q < n; q++) { double phi = slopelimiter(name, r[q], eta[q]); put(fo, &phi, 8); }
  } else if (k == "riemann") {
    int n = rd.get<int32_t>();
    std::string meth = rd.str(8);
    char dir = rd.get<char>();
    double gamma = rd.get<double>();
    const double *wl = rd.dptr((size_t)8 * n), *wr = rd.dptr((size_t)8 * n);
    std::vector<int32_t> st(n, 0);
    std::vector<double> out((size_t)8 * n, 0.0);
    for (int q = 0; q < n; q++) {
      vector<real> vl(wl + 8 * q, wl + 8 * q + 8), vr(wr + 8 * q, wr + 8 * q + 8), vf(8, 0.0);
      try { riemannS(meth, vl, vr, gamma, dir, vf); } catch (...) { st[q] = 1; }
      for (int c = 0; c < 8; c++) out[8 * q + c] = vf[c];
    }
    put(fo, out.data(), 8 * out.size());
    put(fo, st.data(), 4 * st.size());
  } else if (k == "wavespeeds") {
    int n = rd.get<int32_t>();
    int norm = rd.get<int32_t>();
    double gamma = rd.get<double>();
    const double *w = rd.dptr((size_t)8 * n);
    for (int q = 0; q < n; q++) {
      vector<real> v(w + 8 * q, w + 8 * q + 8);
      double o[4];
      wavespeeds(v, gamma, norm, o[0], o[1], o[2], o[3]);
      put(fo, o, 32);
    }
  } else if (k == "recon" || k == "muscl") {
    // header: nblk, NT, step, lim[16], flux[8], gamma, dx
    int nblk = rd.get<int32_t>(), NT = rd.get<int32_t>(), step = rd.get<int32_t>();
    std::string lim = rd.str(16), flux = rd.str(8);
    double gamma = rd.get<double>(), dx = rd.get<double>();
    meshblock dom;
    block_setup(dom, NT, gamma, lim, flux);
    dom.dx[0] = dx; dom.dy[0] = dx;
    for (int b = 0; b < nblk; b++) {
      const double *Uin = rd.dptr((size_t)8 * NT * NT);
      const double *Bin = rd.dptr((size_t)2 * NT * NT);
      int32_t st = 0;
      scratch_zero(dom, NT);
      if (k == "recon") {
        // W given directly; reconstruction only
        tile_in(dom.W, 8, 0, NT, Uin);
        try {
          if (lim == "WENO3") { WENO2D_3rd(dom, "x", 0); WENO2D_3rd(dom, "y", 0); }
          else if (lim == "WENO5") { WENO2D_5th(dom, "x", 0); WENO2D_5th(dom, "y", 0); }
          else celledges(dom, 0, step);
        } catch (...) { st = 1; }
        scratch_out(dom.wxL, NT, fo); scratch_out(dom.wxR, NT, fo);
        scratch_out(dom.wyL, NT, fo); scratch_out(dom.wyR, NT, fo);
      } else {
        for (int i = 0; i < NT; i++) for (int j = 0; j < NT; j++) for (int c = 0; c < 8; c++) {
          dom.ff[i][j][c][0] = 0; dom.gg[i][j][c][0] = 0;
        }
        tile_in(step == 1 ? dom.U : dom.Us, 8, 0, NT, Uin);
        tile_in((real ****)(step == 1 ? dom.Bi : dom.Bis), 2, 0, NT, Bin);
        try { ref_MUSCL2D(dom, 0, step); } catch (...) { st = 1; }
        tile_out(dom.ff, 8, 0, NT, fo); tile_out(dom.gg, 8, 0, NT, fo);
      }
      put(fo, &st, 4);
    }
  } else if (k == "regrid") {
    // header: nops, NT, op(0 refine, 1 coarsen), gamma; per op the 5 tiles
    // (dad + 4 sons) of U[8] and Bi[2] before the op; outputs the 5 tiles after.
    int nops = rd.get<int32_t>(), NT = rd.get<int32_t>(), op = rd.get<int32_t>();
    double gamma = rd.get<double>();
    meshblock dom;
    block_setup(dom, NT, gamma, "MC", "RUSA");
    for (int o = 0; o < nops; o++) {
      for (int s = 0; s < 5; s++) {
        tile_in(dom.U, 8, s, NT, rd.dptr((size_t)8 * NT * NT));
        tile_in((real ****)dom.Bi, 2, s, NT, rd.dptr((size_t)2 * NT * NT));
      }
      int32_t st = 0;
      try {
        if (op == 0) { refine(dom, dom.U, 8, 0, 1, 2, 3, 4); refineBi_CT(dom, (real ****)dom.Bi, 2, 0, 1, 2, 3, 4); }
        else { coarsen(dom, dom.U, 8, 0, 1, 2, 3, 4); coarsenBi_CT(dom, (real ****)dom.Bi, 2, 0, 1, 2, 3, 4); }
      } catch (...) { st = 1; }
      for (int s = 0; s < 5; s++) { tile_out(dom.U, 8, s, NT, fo); tile_out((real ****)dom.Bi, 2, s, NT, fo); }
      put(fo, &st, 4);
    }
  } else if (k == "criteria") {
    int nblk = rd.get<int32_t>(), NT = rd.get<int32_t>();
    meshblock dom;
    block_setup(dom, NT, 1.4, "MC", "RUSA");
    for (int b = 0; b < nblk; b++) {
      int lev = rd.get<int32_t>();
      tile_in(dom.W, 8, 0, NT, rd.dptr((size_t)8 * NT * NT));
      dom.lp[0][0] = lev; dom.iref[0] = false; dom.icoarse[0] = false;
      ref_criteria(dom, 0);
      int32_t o[2] = {dom.iref[0] ? 1 : 0, dom.icoarse[0] ? 1 : 0};
      put(fo, o, 8);
    }
  } else {
    std::fprintf(stderr, "unknown vec kind %s\n", kind);
    return 2;
  }
  std::fclose(fo);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && std::string(argv[1]) == "vec") {
    if (argc != 5) { std::fprintf(stderr, "usage: vec KIND IN OUT\n"); return 2; }
    std::FILE *devnull = std::freopen("/dev/null", "w", stdout);
    (void)devnull;
    return vec_main(argv[2], argv[3], argv[4]);
  }
  if (argc < 2 || std::string(argv[1]) != "run") {
    std::fprintf(stderr, "usage: run --dir D --input 'nx ny nvar|lx ly|cfl|IC lim flux' [--steps K] [--log F] [--dump a,b] [--trace s] [--timing F] [--step-times F] [--randtree seed,pref,pcoarse,rounds]\n");
    return 2;
  }
  std::string input;
  bool quiet = true;
  for (int a = 2; a < argc; a++) {
    std::string s = argv[a];
    auto next = [&]() { return std::string(argv[++a]); };
    if (s == "--dir") g.dir = next();
    else if (s == "--input") input = next();
    else if (s == "--steps") g.steps = std::atoi(next().c_str());
    else if (s == "--log") g.logf = next();
    else if (s == "--dump") { std::stringstream ss(next()); std::string t; while (std::getline(ss, t, ',')) g.dumps.push_back(std::atoi(t.c_str())); }
    else if (s == "--trace") g.trace = std::atoi(next().c_str());
    else if (s == "--timing") g.timingf = next();
    else if (s == "--step-times") g.steptimesf = next();
    else if (s == "--verbose") quiet = false;
    else if (s == "--randtree") {
      std::string v = next();
      std::sscanf(v.c_str(), "%llu,%lf,%lf,%d", &g.seed, &g.p_ref, &g.p_coarse, &g.rand_rounds);
    } else { std::fprintf(stderr, "bad arg %s\n", s.c_str()); return 2; }
  }
  for (char &c : input) if (c == '|') c = '\n';
  mkdir(g.dir.c_str(), 0755);
  if (chdir(g.dir.c_str()) != 0) { std::perror("chdir"); return 3; }
  g.dir = ".";
  { FILE *fp = std::fopen("input.txt", "w"); std::fprintf(fp, "%s\n", input.c_str()); std::fclose(fp); }
  if (!g.logf.empty()) g_log = std::fopen(g.logf.c_str(), "wb");
  if (!g.steptimesf.empty()) g_steptimes = std::fopen(g.steptimesf.c_str(), "wb");
  if (quiet) { std::FILE *r = std::freopen("/dev/null", "w", stdout); (void)r; }
  std::stringstream answers("2\n0\n");
  std::streambuf *old = std::cin.rdbuf(answers.rdbuf());
  int rc = 0;
  try { rc = ref_main(); } catch (StopRun &) { rc = 0; } catch (std::exception &) { rc = 10; }
  std::cin.rdbuf(old);
  auto t1 = std::chrono::steady_clock::now();
  if (g_log) std::fclose(g_log);
  if (g_steptimes) std::fclose(g_steptimes);
  if (!g.timingf.empty()) {
    FILE *fp = std::fopen(g.timingf.c_str(), "w");
    double s = g_t0_set ? std::chrono::duration<double>(t1 - g_t0).count() : 0.0;
    std::fprintf(fp, "{\"loop_seconds\": %.6f, \"steps\": %d, \"rc\": %d}\n", s, g_step + 1, rc);
    std::fclose(fp);
  }
  if (rc == 10) {
    std::fprintf(stderr, "reference threw at step %d\n", g_step);
    if (!g.dumps.empty() && D) dump_state(g.dir + "/state_error.bin", g_t, D->dt);
  }
  return rc;
}
