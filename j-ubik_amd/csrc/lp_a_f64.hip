// double instantiations of the linear-patching kernels for the FFT lengths of
// group A (JUBIK_LP_LENGTHS_A; see linpatch_impl.hpp).
#include "linpatch_impl.hpp"

namespace jubik {
#define JUBIK_INST(l) \
  template int lp_run<double, l>(const LinGeom&, const LinBuffers&, const LinCall&, hipStream_t);
JUBIK_LP_LENGTHS_A(JUBIK_INST)
#undef JUBIK_INST
}  // namespace jubik
