// Column pass of the linear-patching PSF at L = 384 with its DFTs on the
// matrix cores (the split-BF16 DFT-as-matrix transform of an FP32 plan).
//
// A 384-point column DFT is three 128-point DFTs of the decimated
// subsequences x[3 n1 + n2] followed by a radix-3 step,
//   X[k1 + 128 k2] = sum_n2 W_3^(n2 k2) W_384^(n2 k1) Y_n2[k1].
// The 128-point DFTs are real GEMMs on v_mfma_f32_32x32x16_bf16,
//   Re Y = C Xr + S Xi,   Im Y = C Xi - S Xr,
// (C, S the cosine and sine matrices) with each product split into BF16
// halves, a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi, accumulated in FP32; the
// radix-3 step runs in FP32.  The inverse transform is conj(DFT(conj(.)))
// (unnormalised, like the FFT path; the kernel spectra carry the 1 / L^2).
//
// The pass replaces lp_cols<float, 384, COL_FWD / COL_ADJ> outside fold
// mode: the same inputs, outputs, row skipping, kernel-spectrum multiply
// (packed DC/Nyquist column included) and NLL finalisation.  A persistent
// work-group of 8 waves keeps the C and S fragments of rows 0..63 (64 KiB
// of LDS; the symmetry below gives the other rows) and walks the (sample,
// column tile, patch, job) items, 8 columns each, loading the next item
// while it transforms the current one; one LDS area holds a tile as FP32
// columns, as its BF16 halves or as the real product planes.
#include "dft128.hpp"
#include "fft.hpp"
#include "linpatch.hpp"

namespace jubik {

namespace {

constexpr int ML = 384;           // transform length
constexpr int MLH = ML / 2;       // stored half-spectrum columns
constexpr int MV = 8;             // columns per item
constexpr int MLS = ML + 4;       // LDS stride of an FP32 column
constexpr int MNT = 512;          // 8 waves
constexpr int MNTILE = MLH / MV;  // 24 column tiles
constexpr int MVEC = 3 * MV;      // 128-point vectors per tile (column, n2)
constexpr int MCH = MVEC * 16;    // BF16 chunks (8 values) per component
static_assert(MCH <= MNT, "one split chunk per thread");
static_assert(4 * MCH * 16 <= MV * MLS * 8, "the BF16 halves fit the FP32 tile");
static_assert(MV * ML % MNT == 0, "whole tile loads");

// The 128-point DFT matrix F = C - i S is symmetric under r -> 128 - r:
// C[128 - r] = C[r] and S[128 - r] = -S[r].  With the real products on the
// rows r < 64 only,
//   A1 = C xr,  A2 = S xi,  A3 = C xi,  A4 = S xr,
// the outputs are Y[r] = (A1 + A2) + i (A3 - A4) and Y[128 - r] = (A1 - A2)
// + i (A3 + A4) for 0 < r < 64, Y[0] = A1 + i A3 (S row 0 is zero), and
// Y[64] = sum_n (-1)^n x[n], which the split phase sums directly.  This
// halves the matrix-core work and the fragments held in LDS.
constexpr int NFRAG_LDS = 4 * 2 * 8 * 64;  // [C_hi, C_lo, S_hi, S_lo][mt < 2][ks][lane]
constexpr int PST = 71;  // row stride (floats) of the product planes: conflict-free writes
constexpr int PBYTES = 4 * MVEC * PST * 4;
constexpr int BUF_BYTES = (PBYTES > MV * MLS * 8 ? PBYTES : MV * MLS * 8);
static_assert(MV * MLS * 8 <= BUF_BYTES && PBYTES <= BUF_BYTES, "tile area");

// In-place DFT of the MV columns buf[c * MLS + pos], pos < ML: the forward
// transform, or with INV the unnormalised inverse.  F: the row-0..63
// fragments; w1, w2 = W_384^k1 and W_384^(2 k1) for this thread's k1 = tid
// % 128 (the radix-3 tasks of a thread share it); z: MCH + MVEC complex
// values, the chunk sums of Y[64] and their totals.  Begins and ends with every thread's LDS accesses
// ordered by barriers.
template <bool INV>
__device__ void dft384_lds(cplx<float>* buf, const bf16x8* F, cplx<float> w1, cplx<float> w2,
                           cplx<float>* z, int tid) {
  bf16x8* xs = reinterpret_cast<bf16x8*>(buf);
  float* P = reinterpret_cast<float*>(buf);
  // 1. split: chunk t < MCH holds n1 = 8 q + j (j < 8), q = t / MVEC, of
  //    vector vi = t % MVEC = 3 c + n2 (column c, subsequence n2); its four
  //    components (re hi, re lo, im hi, im lo) go to xs[comp * MCH + t] once
  //    every thread has read its FP32 values (the layouts share memory).
  //    The chunk's alternating sum sum_j (-1)^j x[8 q + j] goes to z[t].
  const bool sp = tid < MCH;
  float re[8], im[8];
  if (sp) {
    const int vi = tid % MVEC, q = tid / MVEC;
    const int c = vi / 3, n2 = vi - 3 * c;
    const cplx<float>* col = buf + c * MLS + 24 * q + n2;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const cplx<float> v = col[3 * j];
      re[j] = v.re;
      im[j] = INV ? -v.im : v.im;
    }
  }
  __syncthreads();
  if (sp) {
    bf16x8 rh, rl, ih, il;
    cplx<float> alt = {0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const unsigned short a = bf16_bits(re[j]), b = bf16_bits(re[j] - bf16_value(a));
      const unsigned short c = bf16_bits(im[j]), d = bf16_bits(im[j] - bf16_value(c));
      rh[j] = __builtin_bit_cast(__bf16, a);
      rl[j] = __builtin_bit_cast(__bf16, b);
      ih[j] = __builtin_bit_cast(__bf16, c);
      il[j] = __builtin_bit_cast(__bf16, d);
      alt.re += (j & 1) ? -re[j] : re[j];
      alt.im += (j & 1) ? -im[j] : im[j];
    }
    xs[tid] = rh;
    xs[MCH + tid] = rl;
    xs[2 * MCH + tid] = ih;
    xs[3 * MCH + tid] = il;
    z[tid] = alt;
  }
  __syncthreads();
  // 2. GEMMs: wave wv computes product prod = wv / 2 (A1, A2, A3, A4) on
  //    rows mt*32 .. mt*32+31, mt = wv % 2, for 32 B columns, of which the
  //    first MVEC are the vectors (the rest repeat vectors and are
  //    discarded: an output column depends on its own B column only).  B
  //    lane l holds column l % 32, rows (l / 32) * 8 + 0..7 of the 16-row
  //    step.
  const int lane = tid & 63, wv = tid >> 6;
  const int mt = wv & 1, prod = wv >> 1;
  const int vi = lane & 31, kh = lane >> 5;
  const int vs = vi < MVEC ? vi : vi - MVEC;
  const int aoff = (prod & 1) ? 2 * 2 * 512 : 0;                 // S or C fragments
  const int boff = (prod == 0 || prod == 3) ? 0 : 2 * MCH;       // xr or xi halves
  if (wv == 7 && lane < MVEC) {
    // Y[64] of vector `lane`: the 16 chunk sums in a fixed order
    cplx<float> sum = {0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 16; ++q) sum = cadd(sum, z[q * MVEC + lane]);
    z[MCH + lane] = sum;
  }
  f32x16 acc = {};
#pragma unroll 4
  for (int ks = 0; ks < 8; ++ks) {
    const int xo = boff + (2 * ks + kh) * MVEC + vs;
    const bf16x8 bh = xs[xo], bl = xs[MCH + xo];
    const int o = aoff + (mt * 8 + ks) * 64 + lane;
    const bf16x8 ah = F[o], al = F[2 * 512 + o];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
  }
  __syncthreads();  // every wave has read the BF16 halves
  // product plane P[prod][vi][row], row stride PST; accumulator register r
  // holds row 8 (r / 4) + 4 (lane / 32) + r % 4
  if (vi < MVEC) {
    float* dst = P + (prod * MVEC + vi) * PST + mt * 32 + 4 * kh;
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[8 * (r >> 2) + (r & 3)] = acc[r];
  }
  __syncthreads();
  // 3. radix 3 across the subsequences: task (c, k1) assembles Y_n2[k1]
  //    for n2 = 0..2 from the products (or the alternating sums at k1 = 64)
  //    and writes X[k1], X[k1 + 128], X[k1 + 256]; all reads precede the
  //    writes (the tile overwrites the product planes).
  constexpr int NTASK = MV * 128 / MNT;
  static_assert(MNT % 128 == 0, "one k1 per thread");
  const int k1 = tid & 127;
  const int row = k1 <= 64 ? k1 : 128 - k1;
  const float sg = k1 <= 64 ? 1.f : -1.f;  // sign of A2 and A4
  cplx<float> y[NTASK][3];
#pragma unroll
  for (int h = 0; h < NTASK; ++h) {
    const int c = (tid + h * MNT) >> 7;
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      const int v = 3 * c + n2;
      if (k1 == 64) {
        y[h][n2] = z[MCH + v];
      } else {
        const float* pv = P + v * PST + row;
        const float a1 = pv[0], a2 = pv[MVEC * PST], a3 = pv[2 * MVEC * PST],
                    a4 = pv[3 * MVEC * PST];
        y[h][n2] = cplx<float>{a1 + sg * a2, a3 - sg * a4};
      }
    }
  }
  __syncthreads();
  const float s3 = 0.86602540378443864676f;  // sin(2 pi / 3)
#pragma unroll
  for (int h = 0; h < NTASK; ++h) {
    const int t = tid + h * MNT;
    const cplx<float> t1 = cmul(y[h][1], w1);
    const cplx<float> t2 = cmul(y[h][2], w2);
    // W_3 = exp(-2 pi i / 3): X1 = y0 - (t1 + t2) / 2 - i s3 (t1 - t2),
    // X2 = y0 - (t1 + t2) / 2 + i s3 (t1 - t2)
    const cplx<float> sm = cadd(t1, t2), df = csub(t1, t2);
    const cplx<float> mid = {y[h][0].re - 0.5f * sm.re, y[h][0].im - 0.5f * sm.im};
    const cplx<float> rot = {s3 * df.im, -s3 * df.re};
    cplx<float> x0 = cadd(y[h][0], sm), x1 = cadd(mid, rot), x2 = csub(mid, rot);
    if (INV) {
      x0.im = -x0.im;
      x1.im = -x1.im;
      x2.im = -x2.im;
    }
    cplx<float>* col = buf + (t >> 7) * MLS + k1;
    col[0] = x0;
    col[128] = x1;
    col[256] = x2;
  }
  __syncthreads();
}

template <bool ADJ>
__global__ __launch_bounds__(MNT) void lp_cols_mfma(LinGeom g, const bf16x8* __restrict__ frag,
                                                   const cplx<float>* __restrict__ R1,
                                                   const cplx<float>* __restrict__ KT,
                                                   const cplx<float>* __restrict__ twg,
                                                   cplx<float>* __restrict__ out,
                                                   const cplx<float>* __restrict__ KB,
                                                   NllFinal fin, int nitems,
                                                   const float* __restrict__ tent) {
  constexpr int L2P = half_stride(ML);
  constexpr int PT = MV * ML / MNT;  // tile elements per thread
  __shared__ bf16x8 F[NFRAG_LDS];
  __shared__ __attribute__((aligned(16))) cplx<float> buf[BUF_BYTES / 8];
  __shared__ cplx<float> zs[MCH + MVEC];
  __shared__ double red[MNT / 64];
  const int tid = threadIdx.x;
  // rows 0..63 of each fragment set: the first half of each component
  for (int i = tid; i < NFRAG_LDS; i += MNT) F[i] = frag[(i >> 10) * 2048 + (i & 1023)];
  if constexpr (ADJ) {
    // The first nseg work-groups finalise the NLL of the forward pass that
    // wrote the partials, in the order of lp_cols.
    if (fin.out && (int)blockIdx.x < fin.nseg) {
      const double* pp = fin.partial + (int64_t)blockIdx.x * fin.seglen;
      double acc = 0.0, nb = 0.0;
      for (int64_t k = tid; k < fin.seglen; k += MNT) acc += pp[k];
      acc = block_sum<MNT>(acc, red);
      if (fin.bad_out) {
        const double* pb = pp + (int64_t)fin.nseg * fin.seglen;
        for (int64_t k = tid; k < fin.seglen; k += MNT) nb += pb[k];
        nb = block_sum<MNT>(nb, red);
      }
      if (tid == 0) {
        fin.out[blockIdx.x] = acc;
        if (fin.bad_out) fin.bad_out[blockIdx.x] = nb;
      }
    }
  }
  __syncthreads();
  // Items in rounds of gridDim.x; within a round the work-groups of one XCD
  // take consecutive items, so neighbouring column tiles (one 128-byte line
  // of a row between them) and the samples of a tile share that XCD's L2.
  // The next item's tile and kernel spectrum are loaded into registers
  // while the current one is transformed.
  const cplx<float> w1 = twg[tid & 127], w2 = twg[2 * (tid & 127)];
  const int slot = xcd_tile(blockIdx.x, gridDim.x);
  struct Item {
    int64_t idx, spec, rows;  // output slab, spectrum, shared row spectra of the input
    int k0, row0;
  };
  auto locate = [&](int item) {
    Item it;
    const int smp = item % g.S;
    item /= g.S;
    it.k0 = (item % MNTILE) * MV;
    item /= MNTILE;
    const int p = item % g.P;
    const int z = item / g.P;
    it.idx = (int64_t)(smp * g.J + z) * g.P + p;
    it.rows = ((int64_t)r1_image(g, smp, z, !ADJ) * g.n + p / g.n) * g.H;
    it.spec = spec_index(g, z, p);
    // image row of frame row 0: rows outside the image are zero on input
    // (never stored) and never read on output
    it.row0 = start_x(g, p % g.n) - g.h - g.m;
    return it;
  };
  // element e = tid + u * MNT of a tile: column e % MV, frame row e / MV
  cplx<float> xin[PT], kv[PT], be[PT];
  auto fetch = [&](const Item& it) {
    // the row spectra lp_rows shares between the patches of a patch column
    // (plain loads: neighbouring patches read the same rows); the forward
    // applies the window row's tent weight
    const cplx<float>* in = R1 + it.rows * L2P;
    const cplx<float>* ks = KT + it.spec * L2P * ML;
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * MNT;
      const int c = e & (MV - 1), pos = e >> 3;
      cplx<float> v = {0.f, 0.f};
      if (ADJ) {
        const int ir = it.row0 + pos;
        if (ir >= 0 && ir < g.H) v = in[(int64_t)ir * L2P + it.k0 + c];
      } else {
        const int rr = pos - g.m;
        const int ir = it.row0 + g.m + rr;
        if (rr >= 0 && rr < g.w && ir >= 0 && ir < g.H)
          v = cscale(in[(int64_t)ir * L2P + it.k0 + c], tent[rr]);
      }
      xin[u] = v;
      kv[u] = ks[pos * L2P + it.k0 + c];
      be[u] = (it.k0 == 0 && c == 0) ? KB[it.spec * ML + pos] : cplx<float>{0.f, 0.f};
    }
  };
  if (slot < nitems) fetch(locate(slot));
  for (int base = 0; base + slot < nitems; base += gridDim.x) {
    const Item it = locate(base + slot);
    cplx<float> kc[PT], bc[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * MNT;
      buf[(e & (MV - 1)) * MLS + (e >> 3)] = xin[u];
      kc[u] = kv[u];
      bc[u] = be[u];
    }
    __syncthreads();
    if (base + gridDim.x + slot < nitems) fetch(locate(base + gridDim.x + slot));
    dft384_lds<false>(buf, F, w1, w2, zs, tid);
    // spectrum multiply; the packed column (k = 0) adds beta conj(P[-f]),
    // the conjugated factors for the adjoint
    cplx<float> prod[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * MNT;
      const int c = e & (MV - 1), f = e >> 3;
      const cplx<float> val = buf[c * MLS + f];
      prod[u] = ADJ ? cmulc(val, kc[u]) : cmul(val, kc[u]);
      if (it.k0 == 0 && c == 0) {
        const cplx<float> pm = buf[f == 0 ? 0 : ML - f];
        const cplx<float> pc = {pm.re, -pm.im};
        prod[u] = cadd(prod[u], ADJ ? cmulc(pc, bc[u]) : cmul(pc, bc[u]));
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * MNT;
      buf[(e & (MV - 1)) * MLS + (e >> 3)] = prod[u];
    }
    __syncthreads();
    dft384_lds<true>(buf, F, w1, w2, zs, tid);
    cplx<float>* o = out + it.idx * L2P * (ADJ ? g.w : ML);
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * MNT;
      const int c = e & (MV - 1), pos = e >> 3;
      const cplx<float> v = buf[c * MLS + pos];
      if (ADJ) {
        const int r = pos - g.m;
        if (r >= 0 && r < g.w) st_stream(&o[r * L2P + it.k0 + c], v);
      } else {
        const int ir = it.row0 + pos;
        if (ir >= 0 && ir < g.H) st_stream(&o[pos * L2P + it.k0 + c], v);
      }
    }
    __syncthreads();  // the next item's tile overwrites buf
  }
}

}  // namespace

bool linpatch_mfma_supported(int dtype, const LinGeom& g) {
  return dtype == 0 && g.L == ML && g.Lw == ML;
}

int linpatch_cols_mfma(const LinGeom& g, const LinBuffers& b, int adjoint, NllFinal fin,
                       hipStream_t s) {
  if (!b.frag) return fail(kNotReady, "the plan holds no matrix-core DFT fragments");
  if (g.L != ML || g.Lw != ML) return fail(kUnsupported, "matrix-core column pass needs L = 384");
  const int nitems = g.S * MNTILE * g.P * g.J;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      ncu = v;
  }
  // one work-group per CU (the LDS holds one); the adjoint needs at least
  // nseg of them for the NLL finalisation
  int grid = nitems < ncu ? nitems : ncu;
  if (adjoint && fin.out && grid < fin.nseg) grid = fin.nseg;
  const auto* F = static_cast<const bf16x8*>(b.frag);
  const auto* R1 = static_cast<const cplx<float>*>(b.R1);
  const auto* KT = static_cast<const cplx<float>*>(b.KT);
  const auto* tw = static_cast<const cplx<float>*>(b.tw);
  auto* R2 = static_cast<cplx<float>*>(b.R2);
  const auto* KB = static_cast<const cplx<float>*>(b.KB);
  const auto* tent = static_cast<const float*>(b.tent);
  if (adjoint) {
    TimedLaunch _t("lp_cols_adj_mfma", s);
    lp_cols_mfma<true><<<grid, MNT, 0, s>>>(g, F, R1, KT, tw, R2, KB, fin, nitems, tent);
  } else {
    TimedLaunch _t("lp_cols_fwd_mfma", s);
    lp_cols_mfma<false><<<grid, MNT, 0, s>>>(g, F, R1, KT, tw, R2, KB, NllFinal{}, nitems,
                                              tent);
  }
  return check_hip(hipGetLastError(), "linpatch_cols_mfma");
}

}  // namespace jubik
