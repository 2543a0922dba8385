// Host-side launchers of the stand-alone slicing and convolution primitives.
#pragma once
#include "common.hpp"

namespace jubik {

int slice_patches(int dtype, const void* x, int64_t B, int64_t H, int64_t W, int64_t n,
                  int64_t margin, void* out, hipStream_t s);
int slice_patches_adjoint(int dtype, const void* patches, int64_t B, int64_t H, int64_t W,
                          int64_t n, int64_t margin, void* x, hipStream_t s);
bool convolve_length_supported(int N);
// out (nby, N, N) = dvol * Re(IFFT2(FFT2(x[b / (nby / nbx)]) * FFT2(y[b]))).
// ``tw``: N twiddles; ``work``: nby * N * N complex.
int convolve2d(int dtype, const void* x, const void* y, int64_t nbx, int64_t nby, int64_t N,
               double dvol, const void* tw, void* work, void* out, hipStream_t s);

}  // namespace jubik
