// Linear-patching PSF entry points: dispatch on the element type and the
// FFT length to the per-length instantiations of linpatch_impl.hpp (the
// lp_<group>_f32.hip / lp_<group>_f64.hip translation units).
#include "linpatch.hpp"

namespace jubik {

template <typename T, int L>
int lp_run(const LinGeom& g, const LinBuffers& b, const LinCall& c, hipStream_t s);

#define JUBIK_DECL(l)                                                                      \
  extern template int lp_run<float, l>(const LinGeom&, const LinBuffers&, const LinCall&, \
                                       hipStream_t);                                      \
  extern template int lp_run<double, l>(const LinGeom&, const LinBuffers&, const LinCall&, \
                                        hipStream_t);
JUBIK_LP_LENGTHS(JUBIK_DECL)
#undef JUBIK_DECL

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

int linpatch_fft_length(int w, int m) {
  const int Lw = w + 2 * m;
  if (linpatch_length_supported(Lw)) return Lw;
  // fold mode: the smallest compiled length that holds the window's whole
  // linear convolution (margin >= w / 2) and the Lw window
  const int need = Lw > 2 * w ? Lw : 2 * w;
  int best = 0;
#define JUBIK_CASE(l) \
  if (l >= need && (l - w) % 2 == 0 && (best == 0 || l < best)) best = l;
  JUBIK_LP_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
  return best;
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

static int run(int dtype, const LinGeom& g, const LinBuffers& b, const LinCall& c,
               hipStream_t s) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
  switch (g.L) {
#define JUBIK_CASE(l) \
  case l:             \
    return dtype == 0 ? lp_run<float, l>(g, b, c, s) : lp_run<double, l>(g, b, c, s);
    JUBIK_LP_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    default:
      return fail(kUnsupported, "FFT length " + std::to_string(g.L) + " is not compiled in");
  }
}

int linpatch_prepare(int dtype, const LinGeom& g, const LinBuffers& b, const void* kernel,
                     int layout, hipStream_t s) {
  LinCall c;
  c.op = LinCall::kPrepare;
  c.in = kernel;
  c.layout = layout;
  return run(dtype, g, b, c, s);
}
int linpatch_forward(int dtype, const LinGeom& g, const LinBuffers& b, const void* x, void* y,
                     double scale, hipStream_t s) {
  LinCall c;
  c.op = LinCall::kForward;
  c.in = x;
  c.out = y;
  c.scale = scale;
  return run(dtype, g, b, c, s);
}
int linpatch_forward_poisson(int dtype, const LinGeom& g, const LinBuffers& b, const void* x,
                             double scale, const void* expo, const int* counts, void* resid,
                             void* lam, double* partial, hipStream_t s) {
  LinCall c;
  c.op = LinCall::kForwardPoisson;
  c.in = x;
  c.scale = scale;
  c.expo = expo;
  c.counts = counts;
  c.resid = resid;
  c.lam = lam;
  c.partial = partial;
  return run(dtype, g, b, c, s);
}
int linpatch_forward_metric(int dtype, const LinGeom& g, const LinBuffers& b, const void* dx,
                            double scale, const void* expo, const int* counts, const void* lam,
                            void* cot, hipStream_t s) {
  LinCall c;
  c.op = LinCall::kForwardMetric;
  c.in = dx;
  c.scale = scale;
  c.expo = expo;
  c.counts = counts;
  c.lam = const_cast<void*>(lam);
  c.out = cot;
  return run(dtype, g, b, c, s);
}
int linpatch_adjoint(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy, void* gx,
                     double scale, hipStream_t s) {
  LinCall c;
  c.op = LinCall::kAdjoint;
  c.in = gy;
  c.out = gx;
  c.scale = scale;
  return run(dtype, g, b, c, s);
}
int linpatch_adjoint_gp(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy,
                        double scale, const void* sky, const void* dnu, void* u, hipStream_t s,
                        NllFinal fin, void* ct) {
  LinCall c;
  c.ct = ct;
  c.op = LinCall::kAdjointGp;
  c.in = gy;
  c.scale = scale;
  c.sky = sky;
  c.dnu = dnu;
  c.out = u;
  c.fin = fin;
  return run(dtype, g, b, c, s);
}

}  // namespace jubik
