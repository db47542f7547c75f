// Launch interface of the AMR2D-MHD device kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "amr_types.hpp"

namespace amr {

struct FieldPtrs {
  double* f[kNumFields];
};

// up to four fields filled by one gather/scatter pass
struct FillSet {
  double* base[4];
  int32_t nv[4];
  int32_t kind[4];
  int32_t n;
};

// refine (4 sons from a parent) or coarsen (parent from 4 sons) operation
struct RegridOp {
  int32_t parent_gs;
  int32_t son_gs[4];
  int32_t inst;
  int32_t nb;
  int32_t pad;
};

bool tile_supported(int NT, int ng);

void launch_stage(int NT, int ng, const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip,
                  int step, DevError* err, hipStream_t s, hipEvent_t ev_start = nullptr,
                  hipEvent_t ev_stop = nullptr);
void launch_emf(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, hipStream_t s);
void launch_bupdate(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                    hipStream_t s);
void launch_bsync(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, int step, hipStream_t s);
// ghost fill of up to four fields through a program made of per-instance
// segments (device pointer tables inst_ops/inst_gen, prefix[ninst+1] of op
// counts), gather from the pre-fill state into buf, then scatter
void launch_fill(int NT, const FillOp* const* inst_ops, const FillGen* const* inst_gen, const int32_t* prefix,
                 int ninst, int nops, const FillSet& set, double* buf, hipStream_t s);
// per-leaf CFL bound min(dx,dy)/lambda and refinement indicator; when W is
// non-null it supplies the primitives (initial regrid) instead of U.
void launch_cfl_criteria(int NT, int ng, const FieldPtrs& F, const double* W, const LeafItem* leaves, int nleaf,
                         const InstParams* ip, double* lam_out, double* grad_out, DevError* err, hipStream_t s);
void launch_refine(int NT, int ng, const FieldPtrs& F, const RegridOp* ops, int nops, const InstParams* ip,
                   DevError* err, hipStream_t s);
void launch_coarsen(int NT, int ng, const FieldPtrs& F, const RegridOp* ops, int nops, const int32_t* cmapU,
                    const int32_t* cmapB, const InstParams* ip, DevError* err, hipStream_t s);

// pointwise / tile-level parity entry points
void launch_riemann_batch(int solver, int dir, double gamma, int n, const double* wL, const double* wR, double* flux,
                          int32_t* status, hipStream_t s);
void launch_recon_batch(int NT, int ng, int lim, double dx, int nt, const double* W, double* out, hipStream_t s);

}  // namespace amr
