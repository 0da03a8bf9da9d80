// Adjoint of the pressure projection (placeholder until the adjoint kernels land).
#include "fsl_internal.h"

namespace fsl {

int project_backward(fsl_ctx *c, const fsl_solve_params *, const float *, float *,
                     fsl_solve_stats *) {
  c->err = "adjoint projection not built";
  return FSL_ERR_STATE;
}

}  // namespace fsl
