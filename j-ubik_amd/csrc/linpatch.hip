// Linear-patching PSF entry points: dispatch on the element type to the
// float (linpatch_f32.hip) and double (linpatch_f64.hip) instantiations of
// the kernels in linpatch_impl.hpp.
#include "linpatch.hpp"

namespace jubik {

template <typename T>
int lp_prepare_t(const LinGeom&, const LinBuffers&, const void*, int, hipStream_t);
template <typename T>
int lp_forward_t(const LinGeom&, const LinBuffers&, const void*, void*, double, hipStream_t);
template <typename T>
int lp_forward_poisson_t(const LinGeom&, const LinBuffers&, const void*, double, const void*,
                         const int*, void*, void*, double*, hipStream_t);
template <typename T>
int lp_forward_metric_t(const LinGeom&, const LinBuffers&, const void*, double, const void*,
                        const int*, const void*, void*, hipStream_t);
template <typename T>
int lp_adjoint_t(const LinGeom&, const LinBuffers&, const void*, void*, double, hipStream_t);
template <typename T>
int lp_adjoint_gp_t(const LinGeom&, const LinBuffers&, const void*, double, const void*,
                    const void*, void*, hipStream_t, NllFinal);

bool linpatch_length_supported(int L) {
  switch (L) {
#define JUBIK_CASE(l) case l:
    JUBIK_LP_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    return true;
    default:
      return false;
  }
}

int linpatch_col_tile(int dtype, int L, int col_div) {
  if ((dtype != 0 && dtype != 1) || !linpatch_length_supported(L)) return 0;
  if (col_div != 1 && col_div != 2 && col_div != 4) return 0;
  const int v = tile_vectors(JUBIK_COL_TILE_BYTES, L, dtype == 0 ? 8 : 16) / col_div;
  return (col_div == 1 || v >= 2) ? v : 0;
}

int linpatch_row_tile(int dtype, int L, int row_div) {
  if ((dtype != 0 && dtype != 1) || !linpatch_length_supported(L)) return 0;
  if (row_div != 1 && row_div != 2 && row_div != 4) return 0;
  return tile_vectors(JUBIK_ROW_TILE_BYTES, L, dtype == 0 ? 8 : 16) / row_div;
}

int linpatch_prepare(int dtype, const LinGeom& g, const LinBuffers& b, const void* kernel,
                     int layout, hipStream_t s) {
  if (dtype == 0) return lp_prepare_t<float>(g, b, kernel, layout, s);
  if (dtype == 1) return lp_prepare_t<double>(g, b, kernel, layout, s);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}
int linpatch_forward(int dtype, const LinGeom& g, const LinBuffers& b, const void* x, void* y,
                     double scale, hipStream_t s) {
  if (dtype == 0) return lp_forward_t<float>(g, b, x, y, scale, s);
  if (dtype == 1) return lp_forward_t<double>(g, b, x, y, scale, s);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}
int linpatch_forward_poisson(int dtype, const LinGeom& g, const LinBuffers& b, const void* x,
                             double scale, const void* expo, const int* counts, void* resid,
                             void* lam, double* partial, hipStream_t s) {
  if (dtype == 0)
    return lp_forward_poisson_t<float>(g, b, x, scale, expo, counts, resid, lam, partial, s);
  if (dtype == 1)
    return lp_forward_poisson_t<double>(g, b, x, scale, expo, counts, resid, lam, partial, s);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}
int linpatch_forward_metric(int dtype, const LinGeom& g, const LinBuffers& b, const void* dx,
                            double scale, const void* expo, const int* counts, const void* lam,
                            void* cot, hipStream_t s) {
  if (dtype == 0) return lp_forward_metric_t<float>(g, b, dx, scale, expo, counts, lam, cot, s);
  if (dtype == 1) return lp_forward_metric_t<double>(g, b, dx, scale, expo, counts, lam, cot, s);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}
int linpatch_adjoint(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy, void* gx,
                     double scale, hipStream_t s) {
  if (dtype == 0) return lp_adjoint_t<float>(g, b, gy, gx, scale, s);
  if (dtype == 1) return lp_adjoint_t<double>(g, b, gy, gx, scale, s);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}
int linpatch_adjoint_gp(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy,
                        double scale, const void* sky, const void* dnu, void* u, hipStream_t s,
                        NllFinal fin) {
  if (dtype == 0) return lp_adjoint_gp_t<float>(g, b, gy, scale, sky, dnu, u, s, fin);
  if (dtype == 1) return lp_adjoint_gp_t<double>(g, b, gy, scale, sky, dnu, u, s, fin);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}

}  // namespace jubik
