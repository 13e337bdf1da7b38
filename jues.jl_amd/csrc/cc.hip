// Coupled-cluster drivers (placeholder entry points until the device loop lands).
#include "jues_internal.h"
#include "api_util.h"
namespace jues {
struct CCState {};
void CCDeleter::operator()(CCState* p) const { delete p; }
}  // namespace jues
using namespace jues;
extern "C" {
#define JUES_TODO(ctx) JUES_API_BEGIN(ctx) throw Error(JUES_BAD_ARG, "not implemented"); JUES_API_END(ctx)
jues_status jues_rmp2_grad(jues_ctx* ctx, int, const double*, int, int, const double*, double*, double*, double*) { JUES_TODO(ctx) }
jues_status jues_rccd(jues_ctx* ctx, int, const jues_cc_opts*, const double*, int, const double*, int, int, const double*, double*, double*, jues_iter_cb, void*) { JUES_TODO(ctx) }
jues_status jues_rccsd(jues_ctx* ctx, int, const jues_cc_opts*, const double*, int, const double*, int, int, const double*, double*, double*, double*, jues_iter_cb, void*) { JUES_TODO(ctx) }
jues_status jues_cc_setup(jues_ctx* ctx, int, int, const jues_cc_opts*, const double*, int, const double*, int, int, const double*, const double*, const double*) { JUES_TODO(ctx) }
jues_status jues_cc_iterate(jues_ctx* ctx, int, int, int, jues_iter_info*, int*) { JUES_TODO(ctx) }
jues_status jues_cc_reset_diis(jues_ctx* ctx) { JUES_TODO(ctx) }
jues_status jues_cc_get(jues_ctx* ctx, double*, double*, double*) { JUES_TODO(ctx) }
}
