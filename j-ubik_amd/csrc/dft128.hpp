// Batched 128-point DFTs: FP32 FFT and split-BF16 matrix-core DFT.
#pragma once
#include <vector>

#include "common.hpp"

namespace jubik {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16 with round-to-nearest-even (finite inputs)
__host__ __device__ inline unsigned short bf16_bits(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf16_value(unsigned short h) {
  return __builtin_bit_cast(float, (unsigned)h << 16);
}

constexpr int kDft128TileVectors = 256;
// Fragments of C = cos(2 pi r k / 128) and S = sin(2 pi r k / 128), each
// split into BF16 hi and lo halves: [C_hi, C_lo, S_hi, S_lo][mt < 4][ks <
// 8][lane < 64] x 8 values; a lane's fragment holds row mt*32 + lane%32 and
// columns ks*16 + (lane/32)*8 + 0..7 (the A operand of
// v_mfma_f32_32x32x16_bf16).
void dft128_fragments(std::vector<unsigned short>& out);
int dft128_run(int method, const void* frag, const void* tw, const void* x, void* y, int64_t B,
               int nblocks, hipStream_t s);

}  // namespace jubik
