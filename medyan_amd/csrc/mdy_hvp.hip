// mdy_hvp.hip -- Hessian-vector products (exact-gradient force field).
#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

namespace mdy {

int launch_hvp(Ctx &c, const double *d_x, const double *d_v, double *d_hv, hipStream_t st) {
    (void)c; (void)d_x; (void)d_v; (void)d_hv; (void)st;
    return MDY_ERR_STATE;
}

}  // namespace mdy
