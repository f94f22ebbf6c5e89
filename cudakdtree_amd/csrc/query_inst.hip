// query_inst.hip -- instantiates the batched query kernels for one point
// dimension.  The build compiles this file once per dimension 1..8 with
// -DCUKD_QUERY_D=<D>, so the eight instantiations build in parallel.
#include "query_impl.h"

#ifndef CUKD_QUERY_D
#error "compile with -DCUKD_QUERY_D=<dims>"
#endif

#define CUKD_CAT2(a, b) a##b
#define CUKD_CAT(a, b) CUKD_CAT2(a, b)

namespace cukd_amd {
hipError_t CUKD_CAT(launch_query_, CUKD_QUERY_D)(int method, const QueryArgs &a, hipStream_t s)
{
  return launch_query_d<CUKD_QUERY_D>(method, a, s);
}
}  // namespace cukd_amd
