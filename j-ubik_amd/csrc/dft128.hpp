// Batched 128-point DFTs: FP32 FFT and split-BF16 matrix-core DFT.
#pragma once
#include <vector>

#include "common.hpp"

namespace jubik {

constexpr int kDft128TileVectors = 256;
void dft128_fragments(std::vector<unsigned short>& out);
int dft128_run(int method, const void* frag, const void* tw, const void* x, void* y, int64_t B,
               int nblocks, hipStream_t s);

}  // namespace jubik
