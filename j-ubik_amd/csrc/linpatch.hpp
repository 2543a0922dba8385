// Host-side launchers of the linear-patching PSF kernels.
#pragma once
#include "common.hpp"

namespace jubik {

// Device buffers owned by a linpatch plan (all typed by the plan's dtype).
struct LinBuffers {
  const void* tw;    // L twiddles exp(-2 pi i e / L)
  const void* tent;  // w tent weights (jubik/convolve.py:18-25)
  void* KT;          // prepared spectra, (M, P, Ek) x L x L2p (alpha in column 0)
  void* R1;          // row-pass workspace: shared row spectra (S x J x n x H x L2p) of the
                     // signal and cotangent passes, or (M, P, Ek) x w x L2p of the kernel
                     // preparation; allocated S x J x P x L x L2p (no smaller)
  void* R2;          // column-pass workspace, S x J x P x L x L2p
  void* KB;          // beta of the packed DC/Nyquist column, (M, P, Ek) x L
  int col_div;       // column-pass tile: 1, 2 or 4 (a full, half or quarter tile)
  int row_div;       // row-pass tile: 1, 2 or 4
  // Split-BF16 fragments of the 128-point DFT matrices (dft128_fragments);
  // set: the forward and adjoint column passes run on the matrix cores
  // (FP32 plans at L = 384 outside fold mode, linpatch_mfma_supported).
  const void* frag = nullptr;
};

// FFT lengths compiled into the library, in groups of one translation unit
// per group and element type (lp_<group>_f32.hip, lp_<group>_f64.hip).  A
// window length w + 2m outside the list runs at the smallest listed length
// L >= max(w + 2m, 2w) in fold mode (LinGeom).
#define JUBIK_LP_LENGTHS_A(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128)
#define JUBIK_LP_LENGTHS_B(X) X(192) X(224) X(256)
#define JUBIK_LP_LENGTHS_C(X) X(320) X(324)
#define JUBIK_LP_LENGTHS_D(X) X(384) X(432) X(512)
#define JUBIK_LP_LENGTHS(X) \
  JUBIK_LP_LENGTHS_A(X) JUBIK_LP_LENGTHS_B(X) JUBIK_LP_LENGTHS_C(X) JUBIK_LP_LENGTHS_D(X)

bool linpatch_length_supported(int L);
// FFT length for window w and margin m: w + 2m when compiled in, else the
// fold-mode length; 0 when none is compiled in.
int linpatch_fft_length(int w, int m);

// Columns per work-group of the forward and adjoint column passes for a
// column-tile divisor (1, 2 or 4), or 0 when that divisor is not compiled in
// for L (fewer than 2 columns).
int linpatch_col_tile(int dtype, int L, int col_div);
// Row pairs per work-group of the signal and cotangent row passes for a
// row-tile divisor (1, 2 or 4), or 0 when not compiled in (under one pair).
int linpatch_row_tile(int dtype, int L, int row_div);

// dtype: 0 = float32, 1 = float64.  layout: 0 = full frame (H, W), 1 = cut
// window (w, w).  Kernels are laid out (M, P, Ek, ...).
int linpatch_prepare(int dtype, const LinGeom& g, const LinBuffers& b, const void* kernel,
                     int layout, hipStream_t s);
// y (J, H, W) = PSF(scale * x);  x is (E, H, W) when g.x_shared else (J, H, W).
int linpatch_forward(int dtype, const LinGeom& g, const LinBuffers& b, const void* x, void* y,
                     double scale, hipStream_t s);
// Forward fused with exposure and the Poisson residual: lam = expo * PSF,
// resid = expo * (1 - d / lam) where d >= 0 (else 0), NLL partials per row;
// ``partial`` holds 2 S J H doubles: the row sums of the NLL terms, then the
// row counts of read-out pixels with lam <= 0.
int linpatch_forward_poisson(int dtype, const LinGeom& g, const LinBuffers& b, const void* x,
                             double scale, const void* expo, const int* counts, void* resid,
                             void* lam, double* partial, hipStream_t s);
// Tangent forward for the MGVI metric: cot = expo * (expo * PSF(scale * dx)) /
// lam where d >= 0 (else 0); lam is the forward output at the base point.
int linpatch_forward_metric(int dtype, const LinGeom& g, const LinBuffers& b, const void* dx,
                            double scale, const void* expo, const int* counts, const void* lam,
                            void* cot, hipStream_t s);
// gx = scale * PSF^T(gy); gy (J, H, W); gx (E, H, W) when g.x_shared (summed
// over modules) else (J, H, W).
int linpatch_adjoint(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy, void* gx,
                     double scale, hipStream_t s);
// NLL finalisation folded into the adjoint column pass, which starts after
// the forward has written every partial: out[b] = sum of partial[b * seglen
// .. (b + 1) * seglen) for b < nseg, in a fixed order (no extra launch).
// With bad_out set, bad_out[b] also receives the count of read-out pixels
// with lam <= 0 of segment b (stored nseg * seglen after the NLL terms).
struct NllFinal {
  const double* partial = nullptr;
  int64_t seglen = 0;
  int nseg = 0;
  double* out = nullptr;
  double* bad_out = nullptr;
};

// Adjoint chained into the GP, with v_(s,e) = sum_mod scale * sky[e] *
// PSF^T(gy)_(s, mod, e); optionally finalises the NLL.  Without ct, u gets one
// complex (H, W) plane per sample, sum_e v_(s,e) (1 + i dnu_e) (dnu NULL:
// real part only), read by the GP adjoint with GpGeom::ucomplex.  With ct
// set, u gets the real (S, E, H, W) planes v (the GP adjoint applies dnu) and
// ct receives sum_mod scale * PSF^T(gy)_(s, mod, e), the cotangent of the
// sky the PSF saw (point sources added to the diffuse sky).
int linpatch_adjoint_gp(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy,
                        double scale, const void* sky, const void* dnu, void* u, hipStream_t s,
                        NllFinal fin = NllFinal{}, void* ct = nullptr);

// Matrix-core column pass (lp_mfma.hip): the forward (adjoint = 0) or
// adjoint column pass of an FP32 plan at L = 384 with each column DFT as
// three split-BF16 128-point DFT GEMMs and a radix-3 step; b.frag must be
// set.  Same inputs and outputs as the FFT column pass.
bool linpatch_mfma_supported(int dtype, const LinGeom& g);
int linpatch_cols_mfma(const LinGeom& g, const LinBuffers& b, int adjoint, NllFinal fin,
                       hipStream_t s);

// Arguments of one linpatch operation (the entry points above), passed to
// the per-length instantiations (lp_run<T, L> in linpatch_impl.hpp).
struct LinCall {
  enum Op { kPrepare, kForward, kForwardPoisson, kForwardMetric, kAdjoint, kAdjointGp };
  int op = kForward;
  const void* in = nullptr;  // x, dx, gy or the kernel
  void* out = nullptr;       // y, metric cotangent, gx or the GP cotangent planes
  double scale = 1.0;
  const void* expo = nullptr;
  const int* counts = nullptr;
  void* resid = nullptr;
  void* lam = nullptr;
  double* partial = nullptr;
  const void* sky = nullptr;
  const void* dnu = nullptr;
  int layout = 0;
  NllFinal fin{};
  void* ct = nullptr;        // adjoint into the GP: cotangent of the total sky
};

}  // namespace jubik
