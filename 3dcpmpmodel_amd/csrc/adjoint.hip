// adjoint.hip -- discrete adjoint of the step (no reference counterpart).
#include "internal.h"

using namespace cpmp;

extern "C" int cpmp_adjoint(cpmp_ctx *c, const cpmp_adj_desc *d, double *J, double *g_lo, double *g_hi, double *g_ka) {
    (void)c; (void)d; (void)J; (void)g_lo; (void)g_hi; (void)g_ka;
    return set_error(CPMP_EINVAL, "adjoint not available in this build");
}
