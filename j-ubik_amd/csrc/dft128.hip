// Batched 128-point complex DFTs two ways, for the comparison the 2048^2
// configuration asks for (FP32 FFT vs a split-BF16 DFT on the matrix
// cores):
//   method 0: the library's FP32 Stockham FFT in LDS (two radix stages);
//   method 1: Y = F X as real GEMMs on v_mfma_f32_32x32x16_bf16 with
//             Re Y = C Xr + S Xi, Im Y = C Xi - S Xr (C, S the cosine and
//             sine matrices) and each product split into BF16 halves,
//             a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi, accumulated in FP32.
// The C and S fragments (hi and lo) fill 128 KiB of LDS, loaded once per
// persistent work-group that then streams tiles of 256 vectors.
#include <cstring>
#include <vector>

#include "common.hpp"
#include "dft128.hpp"

namespace jubik {

constexpr int DNT = 512;   // 8 waves, 32 vectors each
constexpr int DVT = 256;   // vectors per tile

struct Split8 {
  bf16x8 hi, lo;
};

__device__ inline Split8 split8(const float* v) {
  Split8 s;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const unsigned short h = bf16_bits(v[j]);
    const unsigned short l = bf16_bits(v[j] - bf16_value(h));
    s.hi[j] = __builtin_bit_cast(__bf16, h);
    s.lo[j] = __builtin_bit_cast(__bf16, l);
  }
  return s;
}

__device__ inline bf16x8 negate8(bf16x8 a) {
#pragma unroll
  for (int j = 0; j < 8; ++j) a[j] = -a[j];
  return a;
}

// frag: [C_hi, C_lo, S_hi, S_lo][mt < 4][ks < 8][lane < 64] bf16x8; a lane's
// fragment holds row mt*32 + lane%32, columns ks*16 + (lane/32)*8 + 0..7.
__global__ __launch_bounds__(DNT) void k_dft128_mfma(const bf16x8* __restrict__ frag,
                                                     const cplx<float>* __restrict__ x,
                                                     cplx<float>* __restrict__ y, int ntile) {
  __shared__ bf16x8 F[4 * 4 * 8 * 64];
  for (int i = threadIdx.x; i < 4 * 4 * 8 * 64; i += DNT) F[i] = frag[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kh = (lane >> 5) * 8;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int64_t n = (int64_t)tile * DVT + wv * 32 + (lane & 31);
    const cplx<float>* xv = x + n * 128;
    Split8 xr[8], xi[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      float re[8], im[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const cplx<float> v = xv[ks * 16 + kh + j];
        re[j] = v.re;
        im[j] = v.im;
      }
      xr[ks] = split8(re);
      xi[ks] = split8(im);
    }
    cplx<float>* yv = y + n * 128;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x16 ar = {}, ai = {};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const int o = (mt * 8 + ks) * 64 + lane;
        const bf16x8 ch = F[0 * 2048 + o], cl = F[1 * 2048 + o];
        const bf16x8 sh = F[2 * 2048 + o], sl = F[3 * 2048 + o];
        ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, xr[ks].hi, ar, 0, 0, 0);
        ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, xr[ks].lo, ar, 0, 0, 0);
        ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cl, xr[ks].hi, ar, 0, 0, 0);
        ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh, xi[ks].hi, ar, 0, 0, 0);
        ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh, xi[ks].lo, ar, 0, 0, 0);
        ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sl, xi[ks].hi, ar, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, xi[ks].hi, ai, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch, xi[ks].lo, ai, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cl, xi[ks].hi, ai, 0, 0, 0);
        const bf16x8 nrh = negate8(xr[ks].hi), nrl = negate8(xr[ks].lo);
        ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh, nrh, ai, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh, nrl, ai, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sl, nrh, ai, 0, 0, 0);
      }
      // accumulator layout: register r holds row 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = mt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        yv[m] = cplx<float>{ar[r], ai[r]};
      }
    }
  }
}

// FP32 FFT of V vectors per work-group.
template <int V>
__global__ __launch_bounds__(256) void k_dft128_fft(const cplx<float>* __restrict__ x,
                                                    cplx<float>* __restrict__ y,
                                                    const cplx<float>* __restrict__ twg) {
  constexpr int L = 128;
  constexpr int VS = lds_stride(L, V);
  __shared__ cplx<float> buf[V * VS];
  __shared__ cplx<float> tw[L];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * V * L;
  load_twiddles(tw, twg, L, tid, 256);
  for (int e = tid; e < V * L; e += 256) buf[(e / L) * VS + (e % L)] = x[base + e];
  __syncthreads();
  fft_lds<float, L, V, VS, 256, false>(buf, tw, tid);
  for (int e = tid; e < V * L; e += 256) y[base + e] = buf[(e / L) * VS + (e % L)];
}

void dft128_fragments(std::vector<unsigned short>& out) {
  out.assign((size_t)4 * 4 * 8 * 64 * 8, 0);
  const long double tau = 6.283185307179586476925286766559005768L;
  for (int mat = 0; mat < 2; ++mat)
    for (int mt = 0; mt < 4; ++mt)
      for (int ks = 0; ks < 8; ++ks)
        for (int lane = 0; lane < 64; ++lane)
          for (int j = 0; j < 8; ++j) {
            const int row = mt * 32 + (lane & 31), col = ks * 16 + (lane >> 5) * 8 + j;
            const long double a = tau * ((row * col) % 128) / 128.0L;
            const float v = (float)(mat == 0 ? cosl(a) : sinl(a));
            const unsigned short h = bf16_bits(v);
            const unsigned short l = bf16_bits(v - bf16_value(h));
            const size_t o = (((size_t)mt * 8 + ks) * 64 + lane) * 8 + j;
            out[(size_t)(2 * mat) * 2048 * 8 + o] = h;
            out[(size_t)(2 * mat + 1) * 2048 * 8 + o] = l;
          }
}

int dft128_run(int method, const void* frag, const void* tw, const void* x, void* y, int64_t B,
               int nblocks, hipStream_t s) {
  if (method == 1) {
    k_dft128_mfma<<<nblocks, DNT, 0, s>>>((const bf16x8*)frag, (const cplx<float>*)x,
                                           (cplx<float>*)y, (int)(B / DVT));
  } else {
    k_dft128_fft<32><<<(unsigned)(B / 32), 256, 0, s>>>((const cplx<float>*)x, (cplx<float>*)y,
                                                          (const cplx<float>*)tw);
  }
  return check_hip(hipGetLastError(), "dft128_run");
}

}  // namespace jubik
