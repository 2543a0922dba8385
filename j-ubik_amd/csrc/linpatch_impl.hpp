// Kernels of the spatially variant PSF by linear patching (jubik/convolve.py:66-145) and
// its adjoint, as three fused passes per direction:
//
//   rows : gather two real window rows per complex vector (tent weights and
//          zero margins applied on the fly), FFT along the row, split the two
//          half spectra and store them row-major (R1[.][r][k], k < L/2).  The
//          real DC and Nyquist terms of a row share slot k = 0 as
//          (X[0], X[L/2]), so a row is exactly L/2 complex values.
//   cols : read V adjacent spectrum columns straight into registers, FFT
//          along the columns, multiply by the prepared kernel spectrum in
//          registers (conjugated for the adjoint), inverse FFT starting from
//          the same registers, store row-major (R2[.][f][k]).  The packed
//          column P = C0 + i CN (DC and Nyquist columns, both real) becomes
//          Q[f] = alpha[f] P[f] + beta[f] conj(P[-f]) with alpha = (K0 + KN)/2,
//          beta = (K0 - KN)/2: the product spectra of both real columns,
//          whose inverse is again the packed pair.
//   out  : per output image row, rebuild the full spectra of two patches per
//          complex vector, inverse FFT, and sum the patch rows that cover the
//          row in a fixed order (gather-form overlap-add, no atomics), with a
//          fused epilogue (plain store, or exposure + Poisson residual + FP64
//          NLL partials, or the chain into the GP adjoint).
//
// Kernel spectra are prepared by the same rows/cols passes from the cut
// window, with the ifftshift applied as (-1)^(k1+k2) and the 1/(sum L^2)
// normalisation folded in (convolve.py:104-131 and 162-176; dvol cancels);
// they are stored KT[spec][f][k] with alpha in column 0, and beta in
// KB[spec][f].
#pragma once
#include <type_traits>

#include "linpatch.hpp"

namespace jubik {

enum RowMode { ROW_SIGNAL = 0, ROW_COTANGENT = 1, ROW_KERNEL_WIN = 2, ROW_KERNEL_FULL = 3 };
enum ColMode { COL_FWD = 0, COL_ADJ = 1, COL_KERNEL = 2 };

constexpr int NT = 256;

// Cotangent window column (in [0, Lw)) of FFT-frame position f in [0, L):
// f itself normally; in fold mode the periodic extension of the Lw window
// (window coordinate f - mf taken modulo Lw into [-m, w + m)).
template <bool FOLD>
__device__ __forceinline__ int cot_col(const LinGeom& g, int f) {
  if constexpr (!FOLD) {
    return f;
  } else {
    int q = (f - g.mf + g.m) % g.Lw;
    return q < 0 ? q + g.Lw : q;
  }
}

// ------------------------------------------------------------------ rows
// Kernel modes: grid (ceil(w/2 / V), P, M*Ek), one spectrum's w window rows
// per (patch, kernel).  Signal and cotangent modes: grid (ceil(H/2 / V), n,
// jobs, or S * E images for a signal shared by the modules (x_shared): its
// row spectra do not depend on the module), the row spectra of image rows
// shared by every patch row: for patch
// column iy and image row i, the row FFT of the window columns of iy
//   signal:    tent(c) * scale * x[i, sy + c - h], c < w, at frame f = c + mf;
//   cotangent: y[i, sy + cot_col(f) - h - m] at frame f < L,
// stored at R1[(job * n + iy) * H + i].  A window's rows then differ from
// these only by the row tent weight (signal), which the column pass applies
// on load: each image row is transformed once per patch column instead of
// once per covering patch (two patch rows forward, L / d in the adjoint).
// VRC row pairs per work-group: RowTile<T, L>::V, or a half or quarter of it
// (chosen by the autotuner for grids with too few row tiles).
// FOLD (cotangent mode): the row of L frame positions f holds the periodic
// extension of the Lw-wide cotangent window (fold mode of LinGeom).
template <typename T, int L, int MODE, int VRC = RowTile<T, L>::V, bool FOLD = false>
__global__ __launch_bounds__(NT) void lp_rows(LinGeom g, const T* __restrict__ src,
                                              const T* __restrict__ tent,
                                              const cplx<T>* __restrict__ twg,
                                              cplx<T>* __restrict__ R1, T scale, int nrows) {
  constexpr int V = VRC;
  constexpr int VS = lds_stride(L, V);
  constexpr int LH = L / 2;
  constexpr int L2P = half_stride(L);
  __shared__ cplx<T> buf[V * VS];
  __shared__ cplx<T> tw[L];
  const int tid = threadIdx.x;
  const int p = blockIdx.y;  // patch (kernel modes) or patch column (signal modes)
  const int z = blockIdx.z;
  const int q0 = blockIdx.x * V;
  const int sy = start_y(g, p < g.n ? p : 0);

  int64_t r1base;
  const T* img;
  if constexpr (MODE == ROW_SIGNAL || MODE == ROW_COTANGENT) {
    // z: the job, or for a shared signal the image (sample, energy) of the
    // sky stack; R1 is laid out by the same index
    r1base = (int64_t)z * g.n + p;
    img = src + (int64_t)z * g.H * g.W;
  } else {
    const int mod = z / g.Ek, ek = z - mod * g.Ek;
    const int64_t spec = ((int64_t)mod * g.P + p) * g.Ek + ek;
    r1base = spec;
    img = src + spec * (MODE == ROW_KERNEL_WIN ? (int64_t)g.w * g.w : (int64_t)g.H * g.W);
  }

  // Every global load (window values, tent weights, twiddles) is issued
  // before the first LDS store, so their latencies overlap.
  constexpr int EPT = (V * L + NT - 1) / NT;
  constexpr int TPT = (L + NT - 1) / NT;
  cplx<T> twr[TPT];
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const int e = tid + u * NT;
    if (e < L) twr[u] = twg[e];
  }
  const double step = tent_step(g.d);
  // Signal rows with L > NT / 2: thread-owned window columns f = tid + u NT
  // and the row pairs of the tile uniform over the work-group, so each
  // column's image index and tent weight is formed once and each row's once
  // (instruction issue, not memory, bounds this prologue).  Shorter rows use
  // the element-wise mapping, which keeps every lane busy.
  constexpr bool kColumns = (MODE == ROW_SIGNAL || MODE == ROW_COTANGENT) && 2 * L > NT;
  const Gather gi = make_gather(img, (int64_t)g.H * g.W);
  if constexpr (kColumns) {
    T av[TPT][V], bv[TPT][V];
    // Column tent weights from the plan's table (the same IEEE double values
    // in T), one zero-reading gather per column.
    const Gather gt = make_gather(tent, MODE == ROW_SIGNAL ? (int64_t)g.w : 0);
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int f = tid + u * NT;
      int j;
      bool cok;
      T wc = T(1);
      if constexpr (MODE == ROW_SIGNAL) {
        const int c = f - g.mf;
        j = sy + c - g.h;
        cok = c >= 0 && c < g.w && j >= 0 && j < g.W;
        wc = gather_load<T>(gt, gather_off<T>(gt, cok, c)) * scale;
      } else {
        j = sy + cot_col<FOLD>(g, f) - g.h - g.m;
        cok = j >= 0 && j < g.W;
      }
      if (L % NT != 0 && f >= L) cok = false;
      // Elements outside the image read zero through the gather, so the
      // loads of all rows are in flight together.
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int i0 = 2 * (q0 + v);  // image row
        const bool ra = cok && i0 < g.H;
        const bool rb = cok && i0 + 1 < g.H;
        av[u][v] = wc * gather_load<T>(gi, gather_off<T>(gi, ra, i0 * g.W + j));
        bv[u][v] = wc * gather_load<T>(gi, gather_off<T>(gi, rb, (i0 + 1) * g.W + j));
      }
    }
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int e = tid + u * NT;
      if (e < L) tw[e] = twr[u];
    }
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int f = tid + u * NT;
      if (L % NT != 0 && f >= L) break;
#pragma unroll
      for (int v = 0; v < V; ++v) buf[v * VS + f] = cplx<T>{av[u][v], bv[u][v]};
    }
  } else {
    T av[EPT], bv[EPT];
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = tid + u * NT;
      av[u] = bv[u] = T(0);
      if (V * L % NT != 0 && e >= V * L) break;
      const int v = e / L, f = e - (e / L) * L;
      const int r0 = 2 * (q0 + v);
      if constexpr (MODE == ROW_SIGNAL || MODE == ROW_COTANGENT) {
        // Zero-reading gathers, as in the column mapping above.
        const int c = f - g.mf;
        const int j = MODE == ROW_SIGNAL ? sy + c - g.h : sy + cot_col<FOLD>(g, f) - g.h - g.m;
        const int i0 = r0;  // image row
        const bool cok = r0 < nrows && (MODE != ROW_SIGNAL || (c >= 0 && c < g.w)) && j >= 0 &&
                         j < g.W;
        const bool ra = cok && i0 < g.H;
        const bool rb = cok && i0 + 1 < g.H;
        const T wc = MODE == ROW_SIGNAL && cok ? tent_weight<T>(g.d, step, c) * scale : T(1);
        av[u] = wc * gather_load<T>(gi, gather_off<T>(gi, ra, i0 * g.W + j));
        bv[u] = wc * gather_load<T>(gi, gather_off<T>(gi, rb, (i0 + 1) * g.W + j));
      } else if (r0 < nrows) {
        if constexpr (MODE == ROW_KERNEL_WIN) {
          const int c = f - g.mf;
          if (c >= 0 && c < g.w) {
            av[u] = img[r0 * g.w + c];
            bv[u] = img[(r0 + 1) * g.w + c];
          }
        } else {  // ROW_KERNEL_FULL: cut[r][c] = K[(r - d) mod H][(c - d) mod W]
          const int c = f - g.mf;
          if (c >= 0 && c < g.w) {
            const int j = (c - g.d + g.W) % g.W;
            const int ia = (r0 - g.d + g.H) % g.H;
            const int ib = (r0 + 1 - g.d + g.H) % g.H;
            av[u] = img[ia * g.W + j];
            bv[u] = img[ib * g.W + j];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int e = tid + u * NT;
      if (e < L) tw[e] = twr[u];
    }
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = tid + u * NT;
      if (V * L % NT != 0 && e >= V * L) break;
      const int v = e / L, f = e - (e / L) * L;
      buf[v * VS + f] = cplx<T>{av[u], bv[u]};
    }
  }
  __syncthreads();
  fft_lds<T, L, V, VS, NT, false, kFewVectorPlan<L, V, NT>>(buf, tw, tid);

  cplx<T>* out = R1 + r1base * L2P * nrows;
  // When L/2 is a multiple of the wave size the row pair v of a store is
  // uniform over each wave, so its row base is scalar.
  constexpr int SPT = (V * LH + NT - 1) / NT;
  constexpr bool kWaveUniform = LH % 64 == 0 && NT % 64 == 0;
#pragma unroll
  for (int u = 0; u < SPT; ++u) {
    const int e = tid + u * NT;
    if (V * LH % NT != 0 && e >= V * LH) break;
    int v = e / LH;
    const int k = e - v * LH;
    if constexpr (kWaveUniform) v = __builtin_amdgcn_readfirstlane(v);
    const int r0 = 2 * (q0 + v);
    if (r0 >= nrows) continue;
    const bool st0 = true;
    const bool st1 = r0 + 1 < nrows;  // an odd image height leaves one row
    const cplx<T> zk = buf[v * VS + k];
    cplx<T> A, B;
    if (k == 0) {
      // X[0] and X[L/2] of both real rows are real: packed (X[0], X[L/2]).
      const cplx<T> zh = buf[v * VS + L / 2];
      A = {zk.re, zh.re};
      B = {zk.im, zh.im};
    } else {
      const cplx<T> zn = buf[v * VS + L - k];
      A = {T(0.5) * (zk.re + zn.re), T(0.5) * (zk.im - zn.im)};
      B = {T(0.5) * (zk.im + zn.im), T(0.5) * (zn.re - zk.re)};
    }
    // Row base (scalar when v is) plus a 32-bit lane byte offset.
    char* orow = reinterpret_cast<char*>(out + (int64_t)r0 * L2P);
    const uint32_t kb = (uint32_t)k * (uint32_t)sizeof(cplx<T>);
    if (st0) st_stream(reinterpret_cast<cplx<T>*>(orow + kb), A);
    if (st1) st_stream(reinterpret_cast<cplx<T>*>(orow + L2P * sizeof(cplx<T>) + kb), B);
  }
}

// ------------------------------------------------------------------ cols
// Forward stages, then (signal modes) the spectrum multiply and the inverse
// stages; the last forward and first inverse stage share registers.
template <typename T, int L, bool INV, int S, int V, int VS, int NT_>
__device__ __forceinline__ void stages_mid(cplx<T>* buf, const cplx<T>* tw, int tid) {
  // stages S .. plan_stages(L) - 2 (all but the last), LDS -> LDS
  if constexpr (S + 1 < plan_stages(L)) {
    using St = Stage<T, L, INV, S, V, VS, NT_>;
    typename St::Regs v;
    St::gather_lds(buf, v, tid);
    St::compute(v, tw, tid);
    __syncthreads();
    St::scatter_lds(buf, v, tid);
    __syncthreads();
    stages_mid<T, L, INV, S + 1, V, VS, NT_>(buf, tw, tid);
  }
}

// Exchange the forward spectrum of the packed column (tile 0, vector 0)
// through LDS so each position can read its mirror P[-f].  Called by the
// whole work-group once every thread has finished reading buf.
template <class FL, typename T, int L>
__device__ __forceinline__ void share_packed(cplx<T>* buf, typename FL::Regs& last, int tid) {
  __syncthreads();
  FL::for_each(last, tid, [&](int vec, int pos, cplx<T>& val) {
    if (vec == 0) buf[pos] = val;
  });
  __syncthreads();
}

// 1-D grid over (sample, column tile, patch, z) with the sample fastest;
// z < M * Ek spectra (kernel mode) or J jobs per sample.  Work-groups are
// remapped so that consecutive items run on one XCD: the S samples of a
// tile follow each other there and read its kernel-spectrum tile from that
// XCD's L2 after the first.
// VC columns per work-group: Tile<T, L>::V, or a half or quarter of it for
// geometries with too few tiles to fill the chip (chosen by the autotuner).
// Every column runs the same butterflies whatever VC is; the instantiations
// differ only in how the compiler contracts multiply-adds, so results agree
// to rounding (measured: at most 2.0e-7 relative in FP32, 2.3e-16 in FP64).
// FOLD: fold mode of LinGeom.  The forward folds the L output rows into the
// Lw window rows (through LDS); the adjoint loads the periodic extension of
// the Lw cotangent rows.
// Signal modes read the shared row spectra of lp_rows (R1[(job * n + iy) *
// H + image row]); the patches of one patch column that share rows (two
// forward, about L / d in the adjoint) run close together on one XCD (the
// grid order below) and meet them in its L2.  The forward multiplies each
// window row by its tent weight on load, through a zero-reading gather so
// that the multiply does not pull each load into a branch of its own.
template <typename T, int L, int MODE, int VC = Tile<T, L>::V, bool FOLD = false>
__global__ __launch_bounds__(NT) void lp_cols(LinGeom g, const cplx<T>* __restrict__ R1,
                                              const cplx<T>* __restrict__ KT,
                                              const cplx<T>* __restrict__ twg,
                                              cplx<T>* __restrict__ out,
                                              cplx<T>* __restrict__ KB, NllFinal fin) {
  constexpr int V = VC;
  constexpr int VS = lds_stride(L, V);
  constexpr int LH = L / 2;
  constexpr int L2P = half_stride(L);
  constexpr int NS = plan_stages(L);
  __shared__ cplx<T> buf[V * VS];
  __shared__ cplx<T> tw[L];
  __shared__ double red[NT / 64];
  __shared__ double total_s;
  const int tid = threadIdx.x;
  constexpr int NTILE = (LH + V - 1) / V;
  const int nsmp = (MODE == COL_KERNEL) ? 1 : g.S;
  int item = xcd_tile(blockIdx.x, gridDim.x);
  const int smp = item % nsmp;
  item /= nsmp;
  const int k0 = (item % NTILE) * V;
  item /= NTILE;
  const int p = item % g.P;
  const int z = item / g.P;

  int64_t idx, spec;
  if constexpr (MODE == COL_KERNEL) {
    const int mod = z / g.Ek, ek = z - mod * g.Ek;
    spec = ((int64_t)mod * g.P + p) * g.Ek + ek;
    idx = spec;
  } else {
    idx = (int64_t)(smp * g.J + z) * g.P + p;
    spec = spec_index(g, z, p);
  }
  const int lw = FOLD ? g.Lw : L;  // window rows of the cotangent / output
  const cplx<T>* in;
  if constexpr (MODE == COL_KERNEL)
    in = R1 + idx * L2P * g.w;
  else
    in = R1 + ((int64_t)r1_image(g, smp, z, MODE == COL_FWD) * g.n + p / g.n) * g.H * L2P;
  const int m = g.mf, w = g.w;

  // Twiddles are fetched into registers now and stored to LDS once the tile
  // loads are issued: stage 0 needs none, and the barrier after it orders
  // the stores before their first use.
  constexpr int TPT = (L + NT - 1) / NT;
  cplx<T> twr[TPT];
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const int e = tid + u * NT;
    if (e < L) twr[u] = twg[e];
  }
  auto store_twiddles = [&]() {
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int e = tid + u * NT;
      if (e < L) tw[e] = twr[u];
    }
  };
  T factor = T(1);
  if constexpr (MODE == COL_KERNEL) {
    double s = 0.0;
    for (int rr = tid; rr < w; rr += NT) s += (double)in[rr * L2P].re;
    s = block_sum<NT>(s, red);
    if (tid == 0) total_s = s;
    __syncthreads();
    factor = T(1.0 / (total_s * (double)L * (double)L));
  }

  // forward stage 0 straight from global memory
  using F0 = Stage<T, L, false, 0, V, VS, NT>;
  using FL = Stage<T, L, false, NS - 1, V, VS, NT>;
  typename FL::Regs last;
  // The kernel spectrum of this tile is loaded first: its latency then
  // overlaps the forward stages instead of stalling the multiply.
  typename FL::Regs kreg;
  if constexpr (MODE != COL_KERNEL) {
    const cplx<T>* ks = KT + spec * L2P * L;
    FL::gather(kreg, tid, [&](int vec, int pos) -> cplx<T> {
      const int k = k0 + vec;
      return k < LH ? ks[pos * L2P + k] : cplx<T>{T(0), T(0)};
    });
  }
  // Image row of window row 0: the adjoint's input rows outside the image
  // are zero and were never stored; the forward's output rows there are
  // never read.
  const int row0 = start_x(g, p % g.n) - g.h - g.m;
  const Gather gin = make_gather(in, MODE == COL_FWD ? (int64_t)g.H * L2P : 0);
  const double tstep = tent_step(g.d);
  auto load_col = [&](int vec, int pos) -> cplx<T> {
    const int k = k0 + vec;
    const cplx<T> zero = {T(0), T(0)};
    if constexpr (MODE == COL_ADJ) {
      if (k >= LH) return zero;
      const int ir = row0 + cot_col<FOLD>(g, pos);
      return (ir >= 0 && ir < g.H) ? ld_stream(&in[ir * L2P + k]) : zero;
    } else if constexpr (MODE == COL_FWD) {
      // the tent weight is applied after all loads are issued (tent_rows):
      // a multiply here would make each task group wait for its loads
      // before the next group's are issued
      const int rr = pos - m;
      const int ir = row0 + g.m + rr;
      const bool ok = k < LH && rr >= 0 && rr < w && ir >= 0 && ir < g.H;
      return gather_load<cplx<T>>(gin, gather_off<cplx<T>>(gin, ok, ir * L2P + k));
    } else {
      if (k >= LH) return zero;
      const int rr = pos - m;
      return (rr >= 0 && rr < w) ? ld_stream(&in[rr * L2P + k]) : zero;
    }
  };
  // forward: window row pos - m carries its tent weight
  auto tent_rows = [&](typename F0::Regs& regs) {
    if constexpr (MODE == COL_FWD)
      F0::for_each(regs, tid, [&](int, int pos, cplx<T>& val) {
        val = cscale(val, tent_weight<T>(g.d, tstep, pos - m));
      });
  };
  static_assert(F0::Ns == 1, "the first stage takes no twiddles");
  if constexpr (NS == 1) {
    F0::gather(last, tid, load_col);
    store_twiddles();
    tent_rows(last);
    __syncthreads();
    F0::compute(last, tw, tid);
  } else {
    typename F0::Regs v0;
    F0::gather(v0, tid, load_col);
    store_twiddles();
    tent_rows(v0);
    F0::compute(v0, tw, tid);
    F0::scatter_lds(buf, v0, tid);
    __syncthreads();
    stages_mid<T, L, false, 1, V, VS, NT>(buf, tw, tid);
    FL::gather_lds(buf, last, tid);
    FL::compute(last, tw, tid);
  }

  if (k0 == 0) share_packed<FL, T, L>(buf, last, tid);
  if constexpr (MODE == COL_KERNEL) {
    cplx<T>* kt = out + spec * L2P * L;
    FL::emit(last, tid, [&](int vec, int pos, cplx<T> val) {
      const int k = k0 + vec;
      if (k >= LH) return;
      if (k == 0) {
        // K0 = (P + conj P-)/2, KN = (P - conj P-)/(2i), each with its
        // ifftshift sign (-1)^(f + k) and the normalisation.
        const cplx<T> pm = buf[pos == 0 ? 0 : L - pos];
        const T s0 = (pos & 1) ? -factor : factor;
        const T sn = ((pos + LH) & 1) ? -factor : factor;
        const cplx<T> k0v = {T(0.5) * s0 * (val.re + pm.re), T(0.5) * s0 * (val.im - pm.im)};
        const cplx<T> knv = {T(0.5) * sn * (val.im + pm.im), T(0.5) * sn * (pm.re - val.re)};
        kt[pos * L2P] = {T(0.5) * (k0v.re + knv.re), T(0.5) * (k0v.im + knv.im)};
        KB[spec * L + pos] = {T(0.5) * (k0v.re - knv.re), T(0.5) * (k0v.im - knv.im)};
      } else {
        const T sgn = ((k + pos) & 1) ? -factor : factor;
        kt[pos * L2P + k] = cscale(val, sgn);
      }
    });
    return;
  } else {
#pragma unroll
    for (int b = 0; b < FL::BPT; ++b)
#pragma unroll
      for (int t = 0; t < FL::R; ++t)
        last[b][t] = (MODE == COL_ADJ) ? cmulc(last[b][t], kreg[b][t]) : cmul(last[b][t], kreg[b][t]);
    if (k0 == 0) {
      // Packed column: alpha P is done above; add beta conj(P[-f]) (the
      // conjugated factor for the adjoint).  buf holds the raw P.
#pragma unroll
      for (int b = 0; b < FL::BPT; ++b) {
        int vec, j;
        if (FL::task(tid, b, vec, j) && vec == 0) {
#pragma unroll
          for (int t = 0; t < FL::R; ++t) {
            const int pos = j + t * FL::NB;
            const cplx<T> pm = buf[pos == 0 ? 0 : L - pos];
            const cplx<T> pc = {pm.re, -pm.im};
            const cplx<T> be = KB[spec * L + pos];
            last[b][t] = cadd(last[b][t], (MODE == COL_ADJ) ? cmulc(pc, be) : cmul(pc, be));
          }
        }
      }
    }
    using I0 = Stage<T, L, true, 0, V, VS, NT>;
    using IL = Stage<T, L, true, NS - 1, V, VS, NT>;
    static_assert(I0::R == FL::R && I0::BPT == FL::BPT, "shared register layout");
    I0::compute(last, tw, tid);
    cplx<T>* o = out + idx * L2P * (MODE == COL_ADJ ? w : lw);
    auto store = [&](int vec, int pos, cplx<T> val) {
      const int k = k0 + vec;
      if (k >= LH) return;
      if constexpr (MODE == COL_ADJ) {
        const int r = pos - m;
        if (r >= 0 && r < w) st_stream(&o[r * L2P + k], val);
      } else {
        // Output rows that fall outside the image are never read by the
        // overlap pass (it gathers image rows only): not stored.
        const int ir = row0 + pos;
        if (ir >= 0 && ir < g.H) st_stream(&o[pos * L2P + k], val);
      }
    };
    // Fold mode, forward: the transformed columns go to LDS, and output row
    // q of the Lw window sums frame positions c, c + Lw and c - Lw, c = q +
    // mf - m (in [0, L) because L >= Lw); rows outside the image are skipped
    // as above.
    auto to_lds = [&](int vec, int pos, cplx<T> val) { buf[vec * VS + pos] = val; };
    auto fold_out = [&]() {
      __syncthreads();
      const int off = g.mf - g.m;
      for (int e = tid; e < V * lw; e += NT) {
        const int vec = e & (V - 1), q = e / V;
        const int k = k0 + vec;
        const int ir = row0 + q;
        if (k >= LH || ir < 0 || ir >= g.H) continue;
        const int c = q + off;
        const cplx<T>* col = buf + vec * VS;
        cplx<T> acc = col[c];
        if (c + lw < L) acc = cadd(acc, col[c + lw]);
        if (c >= lw) acc = cadd(acc, col[c - lw]);
        st_stream(&o[q * L2P + k], acc);
      }
    };
    constexpr bool kFoldOut = FOLD && MODE == COL_FWD;
    if constexpr (NS == 1) {
      if constexpr (kFoldOut) {
        __syncthreads();
        I0::emit(last, tid, to_lds);
        fold_out();
      } else {
        I0::emit(last, tid, store);
      }
    } else {
      __syncthreads();
      I0::scatter_lds(buf, last, tid);
      __syncthreads();
      stages_mid<T, L, true, 1, V, VS, NT>(buf, tw, tid);
      typename IL::Regs v;
      IL::gather_lds(buf, v, tid);
      IL::compute(v, tw, tid);
      if constexpr (kFoldOut) {
        __syncthreads();
        IL::emit(v, tid, to_lds);
        fold_out();
      } else {
        IL::emit(v, tid, store);
      }
    }
  }
  if constexpr (MODE == COL_ADJ) {
    // The first nseg work-groups also finalise the NLL of the forward pass
    // that wrote the partials (same order as a one-block reduction).
    if (fin.out && (int)blockIdx.x < fin.nseg) {
      const double* pp = fin.partial + (int64_t)blockIdx.x * fin.seglen;
      double acc = 0.0, nb = 0.0;
      for (int64_t k = tid; k < fin.seglen; k += NT) acc += pp[k];
      acc = block_sum<NT>(acc, red);
      if (fin.bad_out) {
        const double* pb = pp + (int64_t)fin.nseg * fin.seglen;
        for (int64_t k = tid; k < fin.seglen; k += NT) nb += pb[k];
        nb = block_sum<NT>(nb, red);
      }
      if (tid == 0) {
        fin.out[blockIdx.x] = acc;
        if (fin.bad_out) fin.bad_out[blockIdx.x] = nb;
      }
    }
  }
}

// ------------------------------------------------------------------ overlap
// Threads per work-group of the overlap kernels.
constexpr int NTO = 256;
// The overlap kernels own one output image row per work-group.  Patch rows
// of one patch column (same iy, hence the same column offset) enter the
// row with plain (forward) or tent-weighted (adjoint) sums, which commute
// with the inverse row FFT: their half-spectrum rows are summed first, so
// each work-group runs one inverse FFT per pair of patch columns (patch
// columns iy = 2q, 2q+1 share one complex vector as real and imaginary
// part).  Output columns are owned by threads (j = tid + t * NTO) and summed
// in registers in a fixed order, so results are deterministic.
// Output columns per thread (JPT): W <= JPT * NTO; the host picks 1024, 2048 or
// 4096 columns over NTO threads.

template <typename T, int L>
struct OutTile {
  static constexpr int V = 4;  // iy pairs per pass: n <= 8 in one pass
  static constexpr int LH = L / 2;
  static constexpr int GPT = (V * LH + NTO - 1) / NTO;  // gather tasks (v, kk) per thread
  static constexpr int TPT = (L + NTO - 1) / NTO;       // twiddles per thread
};

// First patch column whose window [s, s + len) can contain canvas column C:
// (C - len) / d + 1 for C >= len.  The quotient comes from the float
// reciprocal rd = 1 / d with one correction step (exact for C < 2^22), in
// place of a full integer division per column.
__device__ __forceinline__ int first_cover(int C, int len, int d, float rd) {
  if (C < len) return 0;
  const int a = C - len;
  int q = (int)((float)a * rd);
  const int r = a - q * d;
  q += (r >= d) - (r < 0);
  return q + 1;
}

// Patch rows covering canvas row R: windows [start_x(ix), start_x(ix) + len)
// with ix in [lo, lo + count).  Starts are strictly increasing, so the set
// is contiguous.  Every argument is uniform over the work-group.
__device__ __forceinline__ void cover_range(const LinGeom& g, int R, int len, int& lo,
                                            int& count) {
  lo = g.n;
  count = 0;
  for (int ix = 0; ix < g.n; ++ix) {
    const int s = start_x(g, ix);
    if (R >= s && R < s + len) {
      lo = min(lo, ix);
      ++count;
    }
  }
}

// buf[v][k] for v < V, k < L: the sum over the covering patch rows ci of
// wt_ci * (A_ci(k) + i B_ci(k)), where A and B are the full spectra of patch
// columns iy = 2(q0 + v) and 2(q0 + v) + 1 rebuilt from their half spectra
// (X(L - kk) = conj X(kk); slot 0 holds the real pair (X[0], X[L/2])).
// Patch (ix, iy) row f starts at base + ((iy n + ix) nrows + f) L2p.  Each
// half-spectrum element is loaded once and fills k = kk and k = L - kk; the
// loads of up to three covering rows are in flight together.  Weights are tent[f] when WEIGHTED.
// Patch offsets (2q + 1) n nrows L2p stay below 2^31 for n <= 64, L <= 384.
template <typename T, int L, bool WEIGHTED>
__device__ __forceinline__ void sum_pair_spectra(cplx<T>* buf, const cplx<T>* __restrict__ base,
                                                 const LinGeom& g, int nrows, int R, int lo,
                                                 int nc, const T* __restrict__ tent, int q0,
                                                 int tid) {
  using OT = OutTile<T, L>;
  constexpr int V = OT::V, LH = OT::LH, GPT = OT::GPT;
  constexpr int VS = lds_stride(L, V);
  constexpr int L2P = half_stride(L);
  // Covering rows per load batch: the forward's windows of L <= 3d rows meet
  // a canvas row at most 3 times in the common geometries, and all three are
  // loaded together.  The adjoint's windows of w = 2d rows meet it twice;
  // they are loaded one after the other, which keeps the chained epilogue
  // at 4 waves per SIMD instead of 3 (measured at 1024^2 x 4E, L = 384:
  // 44 -> 38 us).
  constexpr int CH = WEIGHTED ? 1 : 3;
  const int npair = (g.n + 1) / 2;
  const int pstride = nrows * L2P;  // n <= 64, L <= 384: offsets fit in 31 bits
  int oa[GPT], ob[GPT];
#pragma unroll
  for (int u = 0; u < GPT; ++u) {
    const int e = tid + u * NTO;
    const int v = e / LH, kk = e - (e / LH) * LH;
    const int q = q0 + v;
    const bool ok = (V * LH % NTO == 0 || e < V * LH) && q < npair;
    oa[u] = ok ? 2 * q * g.n * pstride + kk : -1;
    ob[u] = ok && 2 * q + 1 < g.n ? (2 * q + 1) * g.n * pstride + kk : -1;
  }
  cplx<T> A[GPT], B[GPT];
#pragma unroll
  for (int u = 0; u < GPT; ++u) A[u] = B[u] = cplx<T>{T(0), T(0)};
  for (int c0 = 0; c0 < nc; c0 += CH) {
    cplx<T> a[CH][GPT], b[CH][GPT];
    T wt[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      wt[c] = T(0);
#pragma unroll
      for (int u = 0; u < GPT; ++u) a[c][u] = b[c][u] = cplx<T>{T(0), T(0)};
      if (c0 + c < nc) {
        const int ix = lo + c0 + c;
        const int f = R - start_x(g, ix);
        wt[c] = WEIGHTED ? tent[f] : T(1);
        const cplx<T>* rowp = base + ((int64_t)ix * nrows + f) * L2P;
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
          if (oa[u] >= 0) a[c][u] = ld_stream(&rowp[oa[u]]);
          if (ob[u] >= 0) b[c][u] = ld_stream(&rowp[ob[u]]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int u = 0; u < GPT; ++u) {
        if (WEIGHTED) {
          A[u] = cadd(A[u], cscale(a[c][u], wt[c]));
          B[u] = cadd(B[u], cscale(b[c][u], wt[c]));
        } else {
          A[u] = cadd(A[u], a[c][u]);
          B[u] = cadd(B[u], b[c][u]);
        }
      }
  }
#pragma unroll
  for (int u = 0; u < GPT; ++u) {
    const int e = tid + u * NTO;
    if (V * LH % NTO != 0 && e >= V * LH) break;
    const int v = e / LH, kk = e - (e / LH) * LH;
    cplx<T>* dst = buf + v * VS;
    if (kk == 0) {
      dst[0] = cplx<T>{A[u].re, B[u].re};
      dst[LH] = cplx<T>{A[u].im, B[u].im};
    } else {
      dst[kk] = cplx<T>{A[u].re - B[u].im, A[u].im + B[u].re};
      dst[L - kk] = cplx<T>{A[u].re + B[u].im, B[u].re - A[u].im};
    }
  }
}

// grid (H, S * J).  Exposure, counts and (metric, g.sky_shared) the
// base-point lam are per (mod, e) and shared by the samples; y, resid and
// partial are per job, and so is lam when g.sky_shared = 0.
// EPI 1 writes two doubles per (job, row): the NLL terms at partial[job H +
// i] and the number of read-out pixels with lam <= 0 (or NaN) S J H values
// further on.
// FOLD: fold mode of LinGeom; a patch column's output pixel at window
// column c sums frame positions c', c' + Lw and c' - Lw, c' = c - m + mf.
template <typename T, int L, int EPI, int JPT, bool FOLD = false>
__global__ __launch_bounds__(NTO) void lp_fwd_out(LinGeom g, const cplx<T>* __restrict__ R2,
                                                 const cplx<T>* __restrict__ twg,
                                                 T* __restrict__ y, const T* __restrict__ expo,
                                                 const int* __restrict__ counts,
                                                 T* __restrict__ resid,
                                                 double* __restrict__ partial) {
  using OT = OutTile<T, L>;
  constexpr int V = OT::V, TPT = OT::TPT;
  constexpr int VS = lds_stride(L, V);
  constexpr int L2P = half_stride(L);
  __shared__ cplx<T> buf[V * VS];
  __shared__ cplx<T> tw[L];
  __shared__ double red[NTO / 64];
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int job = blockIdx.y;
  const int R = i + g.h + g.m;
  const int lw = FOLD ? g.Lw : L;  // window rows and columns (Lw)
  int lo, nc;
  cover_range(g, R, lw, lo, nc);

  // Twiddles and epilogue operands are fetched up front; their latency
  // overlaps the spectrum loads.
  cplx<T> twr[TPT];
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const int e = tid + u * NTO;
    if (e < L) twr[u] = twg[e];
  }
  T acc[JPT];
#pragma unroll
  for (int t = 0; t < JPT; ++t) acc[t] = T(0);
  const int64_t row = ((int64_t)job * g.H + i) * g.W;
  const int64_t rowd = ((int64_t)(job % g.J) * g.H + i) * g.W;
  // metric: the base-point lam of (mod, e), or of this job's own sample
  // when the samples have their own base points (sky_shared = 0)
  const int64_t rowl = g.sky_shared ? rowd : row;
  const float rd = 1.0f / (float)g.d;
  T ex_v[JPT], lam_v[JPT];
  int cnt_v[JPT];
  const int npair = (g.n + 1) / 2;
  const cplx<T>* base = R2 + (int64_t)job * g.P * lw * L2P;

  // One batch of V patch-column pairs.  The first batch (all of them for
  // n <= 2V) is peeled into straight-line code: inside the batch loop the
  // loads pending at its head (twiddles, epilogue operands) would be waited
  // for before the spectrum loads are issued.
  auto pass = [&](int q0, auto first) {
    // Epilogue operands are fetched ahead of the spectrum loads: they are
    // older in the load queue, so they are in registers by the time the
    // spectra are, and their latency overlaps the spectrum loads instead of
    // the transform (measured at 1024^2 x 4E, L = 384: 53.2 -> 51.2 us).
    if constexpr (decltype(first)::value) {
      if constexpr (EPI != 0) {
#pragma unroll
        for (int t = 0; t < JPT; ++t) {
          const int j = tid + t * NTO;
          ex_v[t] = j < g.W ? expo[rowd + j] : T(0);
          cnt_v[t] = j < g.W ? counts[rowd + j] : -1;
          if (EPI == 2) lam_v[t] = j < g.W ? y[rowl + j] : T(1);
        }
      }
    }
    sum_pair_spectra<T, L, false>(buf, base, g, lw, R, lo, nc, nullptr, q0, tid);
    if constexpr (decltype(first)::value) {
#pragma unroll
      for (int u = 0; u < TPT; ++u) {
        const int e = tid + u * NTO;
        if (e < L) tw[e] = twr[u];
      }
    }
    __syncthreads();
    fft_lds<T, L, V, VS, NTO, true, kFewVectorPlan<L, V, NTO>>(buf, tw, tid);
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) {
        const int C = j + g.h + g.m;
        T s = T(0);
        for (int iy = first_cover(C, lw, g.d, rd); iy < g.n; ++iy) {
          const int sy = start_y(g, iy);
          if (sy > C) break;
          const int c = C - sy;
          const int v = (iy >> 1) - q0;
          if (c < lw && v >= 0 && v < V) {
            if constexpr (FOLD) {
              const int cf = c - g.m + g.mf;
              cplx<T> z = buf[v * VS + cf];
              if (cf + lw < L) z = cadd(z, buf[v * VS + cf + lw]);
              if (cf >= lw) z = cadd(z, buf[v * VS + cf - lw]);
              s += (iy & 1) ? z.im : z.re;
            } else {
              const cplx<T> z = buf[v * VS + c];
              s += (iy & 1) ? z.im : z.re;
            }
          }
        }
        acc[t] += s;
      }
    }
    __syncthreads();
  };
  pass(0, std::true_type{});
  for (int q0 = V; q0 < npair; q0 += V) pass(q0, std::false_type{});

  if constexpr (EPI == 0) {
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) y[row + j] = acc[t];
    }
  } else if constexpr (EPI == 2) {
    // Metric: cotangent expo * (d lam / lam) on read-out pixels, y = lam.
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) {
        const T ex = ex_v[t];
        resid[row + j] = cnt_v[t] >= 0 ? ex * (ex * acc[t]) / lam_v[t] : T(0);
      }
    }
  } else {
    double nll = 0.0, bad = 0.0;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) {
        const T ex = ex_v[t];
        const T lam = ex * acc[t];
        const int d = cnt_v[t];
        T r = T(0);
        if (d >= 0) {
          // A read-out pixel needs lam > 0 (zero exposure must be masked,
          // demos/x-ray-imaging.py:314-315); the terms stay inf or NaN as
          // in the reference and the pixel is counted.
          if (!(lam > T(0))) bad += 1.0;
          // The log of an FP32 lam is taken in FP32 (<= 2 ulp; the gradient
          // does not use it), the terms are summed in FP64.
          const double lg = std::is_same<T, float>::value ? (double)logf((float)lam)
                                                          : log((double)lam);
          nll += (double)lam - (double)d * lg;
          r = ex * (T(1) - T(d) / lam);
        }
        resid[row + j] = r;
        if (y) y[row + j] = lam;
      }
    }
    nll = block_sum<NTO>(nll, red);
    bad = block_sum<NTO>(bad, red);
    if (tid == 0) {
      const int64_t o = (int64_t)job * g.H + i;
      partial[o] = nll;
      partial[o + (int64_t)g.S * g.J * g.H] = bad;
    }
  }
}

// EPI 0: grid (H, nimg); gx[img] = scale * sum over the jobs of img.
// EPI 3: grid (H, S); one complex plane per sample,
//        u[s] = sum_e (sum_mod scale * sky[s, e] * xbar_(s,mod,e)) (1 + i dnu_e)
//        (dnu NULL: real part only), the GP adjoint's column input c with the
//        energies already folded in, jobs summed in their order (s, mod, e).
// EPI 1: grid (H, S * E);  u[s, e] = sum_mod scale * sky[s, e] * xbar_(s,mod,e),
//        one real plane per sample and energy, the cotangent of log s (the GP
//        adjoint folds the energies, with their dnu, while loading them);
//        sky[s, e] is sky[e] when g.sky_shared.
template <typename T, int L, int EPI, int JPT>
__global__ __launch_bounds__(NTO) void lp_adj_out(LinGeom g, const cplx<T>* __restrict__ R2,
                                                 const cplx<T>* __restrict__ twg,
                                                 const T* __restrict__ tent,
                                                 T* __restrict__ gx, T scale,
                                                 const T* __restrict__ sky,
                                                 T* __restrict__ u,
                                                 const T* __restrict__ dnu) {
  using OT = OutTile<T, L>;
  constexpr int V = OT::V, TPT = OT::TPT;
  constexpr int VS = lds_stride(L, V);
  constexpr int L2P = half_stride(L);
  __shared__ cplx<T> buf[V * VS];
  __shared__ cplx<T> tw[L];
  __shared__ T tent_s[L];  // w <= L tent weights
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int Rp = i + g.h;
  int lo, nc;
  cover_range(g, Rp, g.w, lo, nc);

  cplx<T> twr[TPT];
#pragma unroll
  for (int uu = 0; uu < TPT; ++uu) {
    const int e = tid + uu * NTO;
    if (e < L) twr[uu] = twg[e];
  }
  T tnr[TPT];
#pragma unroll
  for (int uu = 0; uu < TPT; ++uu) {
    const int e = tid + uu * NTO;
    if (e < g.w) tnr[uu] = tent[e];
  }
  T acc[JPT], tmp[JPT], tot[JPT];
#pragma unroll
  for (int t = 0; t < JPT; ++t) acc[t] = tot[t] = T(0);
  const int npair = (g.n + 1) / 2;

  // EPI 0 with a shared input and EPI 1: the jobs (s, mod, e) of image
  // blockIdx.y = s * E + e, summed over modules; EPI 0 otherwise: job
  // blockIdx.y.
  const int smp = blockIdx.y / g.E, eb = blockIdx.y - smp * g.E;
  int jb = smp * g.J + eb, je = (smp + 1) * g.J, jstep = g.E;
  if (EPI == 0 && !g.x_shared) {
    jb = blockIdx.y;
    je = blockIdx.y + 1;
    jstep = 1;
  }
  if (EPI == 3) {  // every job of sample blockIdx.y
    jb = blockIdx.y * g.J;
    je = jb + g.J;
    jstep = 1;
  }

  // EPI 1: the sky row of this image, fetched once the first spectrum loads
  // are issued and consumed (its latency then overlaps the transform).
  // EPI 3: the sky row of each job's energy, fetched ahead of its spectrum
  // loads.
  T sky_v[JPT];
  const float rd = 1.0f / (float)g.d;
  // One batch of V patch-column pairs of one job; the first batch of the
  // first job is peeled into straight-line code (see lp_fwd_out).
  auto pass = [&](const cplx<T>* base, int q0, auto first) {
    sum_pair_spectra<T, L, true>(buf, base, g, g.w, Rp, lo, nc, tent, q0, tid);
    if constexpr (decltype(first)::value) {
#pragma unroll
      for (int uu = 0; uu < TPT; ++uu) {
        const int e = tid + uu * NTO;
        if (e < L) tw[e] = twr[uu];
        if (e < g.w) tent_s[e] = tnr[uu];
      }
      if constexpr (EPI == 1) {
        const int simg = g.sky_shared ? eb : blockIdx.y;
        const T* srow = sky + ((int64_t)simg * g.H + i) * g.W;
#pragma unroll
        for (int t = 0; t < JPT; ++t) {
          const int j = tid + t * NTO;
          sky_v[t] = j < g.W ? srow[j] : T(0);
        }
      }
    }
    __syncthreads();
    fft_lds<T, L, V, VS, NTO, true, kFewVectorPlan<L, V, NTO>>(buf, tw, tid);
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) {
        const int Cp = j + g.h;
        T s = T(0);
        for (int iy = first_cover(Cp, g.w, g.d, rd); iy < g.n; ++iy) {
          const int sy = start_y(g, iy);
          if (sy > Cp) break;
          const int c = Cp - sy;
          const int v = (iy >> 1) - q0;
          if (c < g.w && v >= 0 && v < V) {
            const cplx<T> z = buf[v * VS + g.mf + c];
            s += tent_s[c] * ((iy & 1) ? z.im : z.re);
          }
        }
        tmp[t] += s;
      }
    }
    __syncthreads();
  };
  for (int job = jb; job < je; job += jstep) {
    const cplx<T>* base = R2 + (int64_t)job * g.P * g.w * L2P;
    T dn = T(0);
    if constexpr (EPI == 3) {
      const int e = (job - jb) % g.E;
      const int simg = g.sky_shared ? e : blockIdx.y * g.E + e;
      const T* srow = sky + ((int64_t)simg * g.H + i) * g.W;
#pragma unroll
      for (int t = 0; t < JPT; ++t) {
        const int j = tid + t * NTO;
        sky_v[t] = j < g.W ? srow[j] : T(0);
      }
      if (dnu) dn = dnu[e];
    }
#pragma unroll
    for (int t = 0; t < JPT; ++t) tmp[t] = T(0);
    int q0 = 0;
    if (job == jb) {
      pass(base, 0, std::true_type{});
      q0 = V;
    }
    for (; q0 < npair; q0 += V) pass(base, q0, std::false_type{});
    if (EPI == 0) {
#pragma unroll
      for (int t = 0; t < JPT; ++t) acc[t] += tmp[t];
    } else {
#pragma unroll
      for (int t = 0; t < JPT; ++t) {
        const int j = tid + t * NTO;
        if (j < g.W) {
          if (EPI == 3) {
            // folded: real part and dnu-weighted imaginary part (in tot)
            const T v = scale * sky_v[t] * tmp[t];
            acc[t] += v;
            tot[t] += dn * v;
          } else {
            acc[t] += scale * sky_v[t] * tmp[t];
            tot[t] += scale * tmp[t];
          }
        }
      }
    }
  }
  if (EPI == 3) {
    cplx<T>* urow = reinterpret_cast<cplx<T>*>(u) + ((int64_t)blockIdx.y * g.H + i) * g.W;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) urow[j] = cplx<T>{acc[t], tot[t]};
    }
  } else if (EPI == 0) {
    const int64_t row = ((int64_t)blockIdx.y * g.H + i) * g.W;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) gx[row + j] = scale * acc[t];
    }
  } else {
    const int64_t row = ((int64_t)blockIdx.y * g.H + i) * g.W;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NTO;
      if (j < g.W) u[row + j] = acc[t];
    }
    // with point sources: the cotangent of the total sky (u without the
    // diffuse sky factor) into gx
    if (gx) {
#pragma unroll
      for (int t = 0; t < JPT; ++t) {
        const int j = tid + t * NTO;
        if (j < g.W) gx[row + j] = tot[t];
      }
    }
  }
}

}  // namespace jubik

// ====================================================================== host
namespace jubik {

static int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return check_hip(e, what);
  return kOk;
}

template <typename T, int L>
struct LinOps {
  static constexpr int V = Tile<T, L>::V;
  static constexpr int VR = RowTile<T, L>::V;
  static constexpr int LH = L / 2;
  using C = cplx<T>;

  static int prepare(const LinGeom& g, const LinBuffers& b, const void* kernel, int layout,
                     hipStream_t s) {
    dim3 gr((g.w / 2 + VR - 1) / VR, g.P, g.M * g.Ek);
    if (layout == 0)
      {
        TimedLaunch _t("lp_rows_kernel", s);
        lp_rows<T, L, ROW_KERNEL_FULL><<<gr, NT, 0, s>>>(g, (const T*)kernel, (const T*)b.tent,
                                                         (const C*)b.tw, (C*)b.R1, T(1), g.w);
      }
    else
      {
        TimedLaunch _t("lp_rows_kernel", s);
        lp_rows<T, L, ROW_KERNEL_WIN><<<gr, NT, 0, s>>>(g, (const T*)kernel, (const T*)b.tent,
                                                        (const C*)b.tw, (C*)b.R1, T(1), g.w);
      }
    const int gc = (LH + V - 1) / V * g.P * g.M * g.Ek;
    {
      TimedLaunch _t("lp_cols_kernel", s);
      lp_cols<T, L, COL_KERNEL><<<gc, NT, 0, s>>>(g, (const C*)b.R1, nullptr, (const C*)b.tw,
                                                  (C*)b.KT, (C*)b.KB, NllFinal{});
    }
    return launch_status("linpatch_prepare");
  }

  // Fold mode (g.Lw != L) runs the full tiles only.
  template <class F>
  static void with_fold(const LinGeom& g, F&& f) {
    if (g.Lw != L)
      f(std::true_type{});
    else
      f(std::false_type{});
  }

  // The split-BF16 matrix-core column pass replaces the FFT one where the
  // plan holds its fragments (FP32, L = 384, outside fold mode).
  static constexpr bool kMfma = std::is_same<T, float>::value && L == 384;

  static int fwd_front(const LinGeom& g, const LinBuffers& b, const void* x, double scale,
                       hipStream_t s) {
    {
      TimedLaunch _t("lp_rows_fwd", s);
      with_rows(g, b, [&](auto vr) {
        constexpr int VRC = decltype(vr)::value;
        dim3 gr(((g.H + 1) / 2 + VRC - 1) / VRC, g.n, g.S * (g.x_shared ? g.E : g.J));
        lp_rows<T, L, ROW_SIGNAL, VRC><<<gr, NT, 0, s>>>(g, (const T*)x, (const T*)b.tent,
                                                         (const C*)b.tw, (C*)b.R1, T(scale), g.H);
      });
    }
    if constexpr (kMfma)
      if (b.frag && g.Lw == L) return linpatch_cols_mfma(g, b, 0, NllFinal{}, s);
    TimedLaunch _t("lp_cols_fwd", s);
    with_fold(g, [&](auto fold) {
      constexpr bool FOLD = decltype(fold)::value;
      with_cols(g, b, [&](auto vc) {
        constexpr int VC = decltype(vc)::value;
        const int gc = g.S * ((LH + VC - 1) / VC) * g.P * g.J;
        lp_cols<T, L, COL_FWD, VC, FOLD><<<gc, NT, 0, s>>>(g, (const C*)b.R1, (const C*)b.KT,
                                                           (const C*)b.tw, (C*)b.R2, (C*)b.KB,
                                                           NllFinal{});
      });
    });
    return kOk;
  }

  static int adj_front(const LinGeom& g, const LinBuffers& b, const void* gy, hipStream_t s,
                       NllFinal fin = NllFinal{}) {
    int st = kOk;
    with_fold(g, [&](auto fold) {
      constexpr bool FOLD = decltype(fold)::value;
      {
        TimedLaunch _t("lp_rows_adj", s);
        with_rows(g, b, [&](auto vr) {
          constexpr int VRC = decltype(vr)::value;
          dim3 gr(((g.H + 1) / 2 + VRC - 1) / VRC, g.n, g.S * g.J);
          lp_rows<T, L, ROW_COTANGENT, VRC, FOLD><<<gr, NT, 0, s>>>(
              g, (const T*)gy, (const T*)b.tent, (const C*)b.tw, (C*)b.R1, T(1), g.H);
        });
      }
      if constexpr (kMfma && !FOLD) {
        if (b.frag) {
          st = linpatch_cols_mfma(g, b, 1, fin, s);
          return;
        }
      }
      TimedLaunch _t("lp_cols_adj", s);
      with_cols(g, b, [&](auto vc) {
        constexpr int VC = decltype(vc)::value;
        const int gc = g.S * ((LH + VC - 1) / VC) * g.P * g.J;
        lp_cols<T, L, COL_ADJ, VC, FOLD><<<gc, NT, 0, s>>>(g, (const C*)b.R1, (const C*)b.KT,
                                                           (const C*)b.tw, (C*)b.R2, (C*)b.KB,
                                                           fin);
      });
    });
    return st;
  }

  // Row tile of the signal and cotangent row passes: VR / b.row_div row
  // pairs (row_div 1, 2 or 4; at least one pair; fold mode: the full tile).
  template <class F>
  static void with_rows(const LinGeom& g, const LinBuffers& b, F&& f) {
    if (g.Lw == L) {
      if constexpr (VR / 4 >= 1)
        if (b.row_div == 4) return f(std::integral_constant<int, VR / 4>{});
      if constexpr (VR / 2 >= 1)
        if (b.row_div == 2) return f(std::integral_constant<int, VR / 2>{});
    }
    f(std::integral_constant<int, VR>{});
  }

  // Column tile of the forward and adjoint column passes: V / b.col_div
  // columns (col_div 1, 2 or 4; at least 2 columns; fold mode: the full
  // tile).
  template <class F>
  static void with_cols(const LinGeom& g, const LinBuffers& b, F&& f) {
    if (g.Lw == L) {
      if constexpr (V / 4 >= 2)
        if (b.col_div == 4) return f(std::integral_constant<int, V / 4>{});
      if constexpr (V / 2 >= 2)
        if (b.col_div == 2) return f(std::integral_constant<int, V / 2>{});
    }
    f(std::integral_constant<int, V>{});
  }

  // Output columns per thread: W <= 1024, 2048 or 4096 over NTO threads.
  template <class F>
  static void with_jpt(const LinGeom& g, F&& f) {
    if (g.W <= 1024)
      f(std::integral_constant<int, 1024 / NTO>{});
    else if (g.W <= 2048)
      f(std::integral_constant<int, 2048 / NTO>{});
    else
      f(std::integral_constant<int, 4096 / NTO>{});
  }

  template <int EPI>
  static void fwd_out(const LinGeom& g, const LinBuffers& b, T* y, const T* expo,
                      const int* counts, T* resid, double* partial, hipStream_t s,
                      const char* name) {
    TimedLaunch _t(name, s);
    with_fold(g, [&](auto fold) {
      constexpr bool FOLD = decltype(fold)::value;
      with_jpt(g, [&](auto jpt) {
        lp_fwd_out<T, L, EPI, decltype(jpt)::value, FOLD><<<dim3(g.H, g.S * g.J), NTO, 0, s>>>(
            g, (const C*)b.R2, (const C*)b.tw, y, expo, counts, resid, partial);
      });
    });
  }

  template <int EPI>
  static void adj_out(const LinGeom& g, const LinBuffers& b, int nimg, T* gx, double scale,
                      const T* sky, T* u, hipStream_t s, const char* name,
                      const T* dnu = nullptr) {
    TimedLaunch _t(name, s);
    with_jpt(g, [&](auto jpt) {
      lp_adj_out<T, L, EPI, decltype(jpt)::value><<<dim3(g.H, nimg), NTO, 0, s>>>(
          g, (const C*)b.R2, (const C*)b.tw, (const T*)b.tent, gx, T(scale), sky, u, dnu);
    });
  }

  static int run(const LinGeom& g, const LinBuffers& b, const LinCall& c, hipStream_t s) {
    switch (c.op) {
      case LinCall::kPrepare:
        return prepare(g, b, c.in, c.layout, s);
      case LinCall::kForward:
        if (int st = fwd_front(g, b, c.in, c.scale, s)) return st;
        fwd_out<0>(g, b, (T*)c.out, nullptr, nullptr, nullptr, nullptr, s, "lp_fwd_out");
        return launch_status("linpatch_forward");
      case LinCall::kForwardPoisson:
        if (int st = fwd_front(g, b, c.in, c.scale, s)) return st;
        fwd_out<1>(g, b, (T*)c.lam, (const T*)c.expo, c.counts, (T*)c.resid, c.partial, s,
                   "lp_fwd_out_poisson");
        return launch_status("linpatch_forward_poisson");
      case LinCall::kForwardMetric:
        if (int st = fwd_front(g, b, c.in, c.scale, s)) return st;
        fwd_out<2>(g, b, (T*)c.lam, (const T*)c.expo, c.counts, (T*)c.out, nullptr, s,
                   "lp_fwd_out_metric");
        return launch_status("linpatch_forward_metric");
      case LinCall::kAdjoint:
        if (int st = adj_front(g, b, c.in, s)) return st;
        adj_out<0>(g, b, g.S * (g.x_shared ? g.E : g.J), (T*)c.out, c.scale, nullptr, nullptr,
                   s, "lp_adj_out");
        return launch_status("linpatch_adjoint");
      case LinCall::kAdjointGp:
        if (int st = adj_front(g, b, c.in, s, c.fin)) return st;
        // without point sources: one folded complex plane per sample;
        // with them: a real plane per sample and energy, and the total-sky
        // cotangent in c.ct
        if (c.ct)
          adj_out<1>(g, b, g.S * g.E, (T*)c.ct, c.scale, (const T*)c.sky, (T*)c.out, s,
                     "lp_adj_out_gp");
        else
          adj_out<3>(g, b, g.S, nullptr, c.scale, (const T*)c.sky, (T*)c.out, s,
                     "lp_adj_out_gp", (const T*)c.dnu);
        return launch_status("linpatch_adjoint_gp");
      default:
        return fail(kInvalidArgument, "unknown linpatch operation");
    }
  }
};

// One entry per (element type, FFT length); the lp_*.hip translation units
// instantiate them in groups of lengths, linpatch.hip dispatches.
template <typename T, int L>
int lp_run(const LinGeom& g, const LinBuffers& b, const LinCall& c, hipStream_t s) {
  return LinOps<T, L>::run(g, b, c, s);
}

}  // namespace jubik
