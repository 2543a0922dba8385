// float instantiation of the linear-patching kernels (see linpatch_impl.hpp).
#include "linpatch_impl.hpp"

namespace jubik {
template int lp_prepare_t<float>(const LinGeom&, const LinBuffers&, const void*, int, hipStream_t);
template int lp_forward_t<float>(const LinGeom&, const LinBuffers&, const void*, void*, double,
                               hipStream_t);
template int lp_forward_poisson_t<float>(const LinGeom&, const LinBuffers&, const void*, double,
                                       const void*, const int*, void*, void*, double*,
                                       hipStream_t);
template int lp_forward_metric_t<float>(const LinGeom&, const LinBuffers&, const void*, double,
                                      const void*, const int*, const void*, void*, hipStream_t);
template int lp_adjoint_t<float>(const LinGeom&, const LinBuffers&, const void*, void*, double,
                               hipStream_t);
template int lp_adjoint_gp_t<float>(const LinGeom&, const LinBuffers&, const void*, double,
                                  const void*, const void*, void*, hipStream_t, NllFinal);
}  // namespace jubik
