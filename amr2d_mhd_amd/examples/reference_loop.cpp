// The reference's time loop (main.cpp:97-247) driven through the C ABI of
// include/amr2d_mhd.h: what a maintainer of the reference writes in place of
// the body of main() once the hot path lives in libamr2d_mhd.so.  The
// variable names are the reference's (input.txt fields main.cpp:43-69,
// nghosts/maxlevs defs.hpp:27-28); save4py's periodic output becomes an
// amr_download of the slot fields.
//
//   reference_loop IC limiter flux maxlevs [nghosts [max_steps [nx [CFL]]]]
//
// prints one line: steps, t, dt, last_active, nleaf and an FNV-1a hash of U
// over the active slots, so a run can be compared with another binding.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "amr2d_mhd.h"

using std::exception;

namespace {

uint64_t fnv1a(const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  uint64_t h = 1469598103934665603ull;
  for (size_t k = 0; k < n; k++) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}

void report(amr_ctx* ctx, int rc) {
  amr_diag d{};
  amr_last_error(ctx, &d);
  std::fprintf(stderr, "amr2d_mhd error %d at instance %d block %d cell (%d, %d)\n", rc, d.inst, d.nb, d.i, d.j);
}

int run_like_reference_main(int nx, int nghosts, int maxlevs, double lenx, double CFL, const std::string& IC,
                            const std::string& limiter, const std::string& fluxMth, int max_steps) {
  amr_sim_config cfg = {};
  cfg.nx = nx;
  cfg.nghosts = nghosts;
  cfg.maxlevs = maxlevs;
  cfg.lenx = lenx;
  cfg.cfl = CFL;
  std::strncpy(cfg.ic, IC.c_str(), sizeof(cfg.ic) - 1);
  std::strncpy(cfg.limiter, limiter.c_str(), sizeof(cfg.limiter) - 1);
  std::strncpy(cfg.flux, fluxMth.c_str(), sizeof(cfg.flux) - 1);
  amr_ctx* ctx = nullptr;
  if (int rc = amr_create(1, &cfg, 0, &ctx)) {
    std::fprintf(stderr, "amr_create failed: %d\n", rc);
    return rc;
  }
  if (int rc = amr_init(ctx)) {
    report(ctx, rc);
    amr_destroy(ctx);
    return rc;
  }
  int active = 1, steps = 0;
  amr_inst_info info{};
  while (active && (max_steps < 0 || steps < max_steps)) {
    if (int rc = amr_step(ctx, &active)) {  // where the reference prints and throws
      report(ctx, rc);
      amr_destroy(ctx);
      return rc;
    }
    steps++;
    amr_info(ctx, 0, &info);
  }
  amr_info(ctx, 0, &info);
  // save4py's data, from the device: U of every active slot
  const size_t NT = (size_t)nx + 2 * (size_t)nghosts;
  std::vector<double> U((size_t)info.last_active * 8 * NT * NT);
  if (int rc = amr_download(ctx, 0, AMR_U, info.last_active, U.data())) {
    report(ctx, rc);
    amr_destroy(ctx);
    return rc;
  }
  std::printf("steps %d t %.17g dt %.17g last_active %d nleaf %d hashU %llu\n", steps, info.t, info.dt,
              info.last_active, info.nleaf, (unsigned long long)fnv1a(U.data(), U.size() * sizeof(double)));
  amr_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s IC limiter flux maxlevs [nghosts [max_steps [nx [CFL]]]]\n", argv[0]);
    return 2;
  }
  const int maxlevs = std::atoi(argv[4]);
  const int nghosts = argc > 5 ? std::atoi(argv[5]) : 2;
  const int max_steps = argc > 6 ? std::atoi(argv[6]) : -1;
  const int nx = argc > 7 ? std::atoi(argv[7]) : 16;
  const double CFL = argc > 8 ? std::atof(argv[8]) : 0.5;
  try {
    return run_like_reference_main(nx, nghosts, maxlevs, 1.0, CFL, argv[1], argv[2], argv[3], max_steps);
  } catch (const exception&) {
    return 3;
  }
}
