// Host-side launchers of the stand-alone slicing and convolution primitives.
#pragma once
#include "common.hpp"

namespace jubik {

int slice_patches(int dtype, const void* x, int64_t B, int64_t H, int64_t W, int64_t n,
                  int64_t margin, void* out, hipStream_t s);
int slice_patches_adjoint(int dtype, const void* patches, int64_t B, int64_t H, int64_t W,
                          int64_t n, int64_t margin, void* x, hipStream_t s);
// FFT length of a convolve axis of length n: n when compiled in, else the
// smallest compiled length >= 2 n - 1 (zero-padded and folded); 0 when none.
int convolve_fft_length(int64_t n);
// out (nby, H, W) = dvol * Re(IFFT2(FFT2(x[b / (nby / nbx)]) * FFT2(y[b]))),
// the circular convolution over the two axes.  ``twh``, ``tww``: the LH and
// LW twiddles of the axes' FFT lengths; ``work``: nby * LH * W complex.
int convolve2d(int dtype, const void* x, const void* y, int64_t nbx, int64_t nby, int64_t H,
               int64_t W, double dvol, const void* twh, const void* tww, void* work, void* out,
               hipStream_t s);

}  // namespace jubik
