// Operation count of the stage arithmetic on the host (flop_count.cpp).
#pragma once

namespace amr {

// +/- (add), * (mul), / (div) and sqrt performed
struct StageFlops {
  double add = 0., mul = 0., div = 0., sqrt = 0.;
};

// Adds to acc the operations one RK2 stage (step 1 or 2) of one NT x NT leaf
// tile performs on its data: Qin = the stage input [8][NT][NT], B = its face
// field [2][NT][NT], limiter and solver ids as in amr_types.hpp.
void count_stage_tile(int NT, int NG, const double* Qin, const double* B, double gamma, double dxl, double dt,
                      int lim, int solver, int step, StageFlops& acc);

}  // namespace amr
