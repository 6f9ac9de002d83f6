// fr_mlp.hip -- learned troubled-cell sensor (placeholder until the MLP kernels land).
#include "fr2d_internal.h"

namespace fr {

int fr2d_launch_mlp_sensor(const Fr2dCtx &c, const double *d_u, hipStream_t st) {
  (void)c;
  (void)d_u;
  (void)st;
  return fail(FR_E_UNSUPPORTED, "learned sensor not available");
}

}  // namespace fr
