// MP2 adjoint (placeholder until the back-transform lands).
#include "jues_internal.h"
#include "api_util.h"
using namespace jues;
extern "C" jues_status jues_rmp2_grad(jues_ctx* ctx, int, const double*, int, int, const double*, double*, double*,
                                      double*) {
  JUES_API_BEGIN(ctx)
  throw Error(JUES_BAD_ARG, "not implemented");
  JUES_API_END(ctx)
}
