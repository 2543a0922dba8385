// double instantiations of the linear-patching kernels for the FFT lengths of
// group C (JUBIK_LP_LENGTHS_C; see linpatch_impl.hpp).
#include "linpatch_impl.hpp"

namespace jubik {
#define JUBIK_INST(l) \
  template int lp_run<double, l>(const LinGeom&, const LinBuffers&, const LinCall&, hipStream_t);
JUBIK_LP_LENGTHS_C(JUBIK_INST)
#undef JUBIK_INST
}  // namespace jubik
