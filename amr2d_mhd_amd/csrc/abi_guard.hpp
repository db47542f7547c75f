// Exception barrier of the C ABI: every entry point returns a status code and
// no C++ exception crosses the interface (a failed host allocation inside an
// entry point, the only exception this code can raise, becomes AMR_ERR_HOST).
#pragma once
#include <new>

#include "../../include/amr2d_mhd.h"

#define AMR_TRY try {
#define AMR_CATCH                    \
  }                                  \
  catch (const std::bad_alloc&) {    \
    return AMR_ERR_HOST;             \
  }                                  \
  catch (...) {                      \
    return AMR_ERR_HOST;             \
  }
// for the entry points that return a measurement (negative: failure)
#define AMR_CATCH_NEG                \
  }                                  \
  catch (...) {                      \
    return -1;                       \
  }
