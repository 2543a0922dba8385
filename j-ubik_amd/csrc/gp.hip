// Correlated-field GP sky on an N x N grid (fixed amplitude field):
//
//   log s_e = zm + HT(amp_ref * xi_ref) + (HT(amp_alpha * xi_alpha) + mu) * dnu_e
//
// (jubik/sky_model/multifrequency/spectral_product_mf_sky.py:171-219,281-298;
// HT(u) = Re(FFT u) + Im(FFT u), nifty.re's hartley).  Both real transforms
// share one complex 2-D FFT of c = amp_ref xi_ref + i amp_alpha xi_alpha:
// with Z = FFT(c) and Z' = Z(-k),
//   HT(a) = (Z.re + Z'.re + Z.im - Z'.im) / 2,
//   HT(b) = (Z.im + Z'.im - Z.re + Z'.re) / 2.
// Pass 1 transforms columns (a tile of adjacent columns per work-group);
// pass 2 transforms the row pair (k, -k) in one work-group, so the pairing
// needs no extra pass, and applies the epilogue (exp + crop to the sky grid,
// or the amplitude multiply of the adjoint, or the tangent map).
#include "gp.hpp"

namespace jubik {

constexpr int GNT = 256;
// The row pass runs 512 threads: its N/2 + 1 work-groups then hold twice the
// waves per CU, which hides more of its load -> transform -> exp chain
// (measured at N = 1024: forward 15.4 -> 13.0 us; the column pass is faster
// at 256).
constexpr int GNR = 512;

// Columns per work-group of the column pass: the LDS tile width, at most N.
// The adjoint (MODE 1) gathers uplanes cotangent planes per column and runs
// at half the width: twice the work-groups hide that gather's latency
// better (measured at N = 1024: 30 -> 23 us per launch).
template <typename T, int N, int MODE = 0>
struct GpTile {
  static constexpr int V0 = Tile<T, N>::V < N ? Tile<T, N>::V : N;
  static constexpr int V = (MODE == 1 && V0 > 2) ? V0 / 2 : V0;
};

// grid (ceil(N / V), S).  MODE 0: c = amp * xi (forward / tangent);  MODE 1:
// c = sum_p u_p * (1 + i dnu[p mod E]) over the uplanes real (H, W)
// cotangent planes u_p (one per energy, or per sample and energy), zero
// outside the sky grid; with g.ucomplex, c = sum_p u_p over uplanes complex
// planes that hold the energies folded in already.  Sample blockIdx.y reads xi[s] / u[s] and writes
// work[s].
// When V does not divide N (odd grids such as 567) the last tile's columns
// beyond N read zero and are not stored.
template <typename T, int N, int MODE>
__global__ __launch_bounds__(GNT) void gp_cols(GpGeom g, const T* __restrict__ amp_ref,
                                               const T* __restrict__ amp_alpha,
                                               const T* __restrict__ xi_ref,
                                               const T* __restrict__ xi_alpha,
                                               const T* __restrict__ u,
                                               const T* __restrict__ dnu,
                                               const cplx<T>* __restrict__ twg,
                                               cplx<T>* __restrict__ Tw) {
  constexpr int V = GpTile<T, N, MODE>::V;
  constexpr int VS = lds_stride(N, V);
  __shared__ cplx<T> buf[V * VS];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  const int c0 = xcd_tile(blockIdx.x, gridDim.x) * V;
  const int64_t smp = blockIdx.y;
  if (xi_ref) xi_ref += smp * N * N;
  if (xi_alpha) xi_alpha += smp * N * N;
  if (u) u += smp * g.uplanes * g.H * g.W * (g.ucomplex ? 2 : 1);
  Tw += smp * N * N;
  // Loads are issued in unrolled chunks of CH elements per thread (their
  // latencies overlap), twiddles first.
  constexpr int TPT = (N + GNT - 1) / GNT;
  cplx<T> twr[TPT];
#pragma unroll
  for (int u2 = 0; u2 < TPT; ++u2) {
    const int e = tid + u2 * GNT;
    if (e < N) twr[u2] = twg[e];
  }
  constexpr int EPT = (V * N + GNT - 1) / GNT;
  constexpr int CH = (MODE == 0 ? 16 : 8) / (sizeof(T) / 4);
  constexpr int CHK = EPT < CH ? EPT : CH;
  // MODE 1 reads the planes through zero-reading gathers (elements outside
  // the sky grid read zero), one buffer resource per plane (so offsets stay
  // 32-bit however many planes), planes outer and the chunk unrolled inside,
  // so a chunk's loads of one plane are in flight together.
  const int64_t plane = (int64_t)g.H * g.W;
  const int64_t nn = MODE == 0 ? (int64_t)N * N : 0;
  const Gather gar = make_gather(amp_ref, nn), gxr = make_gather(xi_ref, nn);
  const Gather gaa = make_gather(amp_alpha, g.has_alpha ? nn : 0);
  const Gather gxa = make_gather(xi_alpha, g.has_alpha ? nn : 0);
  // Row-vector gathers: the tile's V adjacent columns of a row are one 8- or
  // 16-byte load per array, V times fewer loads than one per element (the
  // adjoint's planes need W to be a multiple of V, so that a tile lies wholly
  // inside or wholly outside the sky grid).  Same products and sums per
  // element as the per-element path below.
  constexpr bool kRowVec = (V * sizeof(T) == 8 || V * sizeof(T) == 16) && N % V == 0;
  // complex planes: a row of the tile is V complex values
  constexpr bool kRowVecC = 2 * V * sizeof(T) == 16 && N % V == 0;
  bool row_vec = false;
  if constexpr (kRowVec) {
    row_vec = MODE == 0 || (g.W % V == 0 && (!g.ucomplex || kRowVecC));
    if (row_vec) {
      using VT = T __attribute__((ext_vector_type(V)));
      constexpr int RPT = (N + GNT - 1) / GNT;  // rows per thread
      constexpr int CR = RPT < 4 ? RPT : 4;     // rows per chunk
      for (int r0 = 0; r0 < RPT; r0 += CR) {
        VT va[CR], vb[CR];
#pragma unroll
        for (int c = 0; c < CR; ++c) {
          const int x1 = tid + (r0 + c) * GNT;
          if constexpr (MODE == 0) {
            const bool ok = x1 < N;
            const int o = x1 * N + c0;
            const int orf = gather_off<T>(gar, ok, o);
            const int oal = gather_off<T>(gaa, ok && g.has_alpha, o);
            va[c] = gather_load<VT>(gar, orf) * gather_load<VT>(gxr, orf);
            vb[c] = gather_load<VT>(gaa, oal) * gather_load<VT>(gxa, oal);
          } else {
            va[c] = VT(T(0));
            vb[c] = VT(T(0));
          }
        }
        if constexpr (MODE == 1) {
          if (g.ucomplex) {
            if constexpr (kRowVecC) {
              using VT2 = T __attribute__((ext_vector_type(2 * V)));
              for (int pl = 0; pl < g.uplanes; ++pl) {
                const cplx<T>* up = reinterpret_cast<const cplx<T>*>(u) + pl * plane;
                const Gather gu = make_gather(up, plane);
#pragma unroll
                for (int c = 0; c < CR; ++c) {
                  const int x1 = tid + (r0 + c) * GNT;
                  const bool ok = x1 < g.H && c0 < g.W;
                  const VT2 w =
                      gather_load<VT2>(gu, gather_off<cplx<T>>(gu, ok, x1 * g.W + c0));
#pragma unroll
                  for (int v = 0; v < V; ++v) {
                    va[c][v] += w[2 * v];
                    vb[c][v] += w[2 * v + 1];
                  }
                }
              }
            }
          } else {
            for (int pl = 0; pl < g.uplanes; ++pl) {
              const Gather gu = make_gather(u + pl * plane, plane);
              const T dn = g.has_alpha ? dnu[pl % g.E] : T(0);
#pragma unroll
              for (int c = 0; c < CR; ++c) {
                const int x1 = tid + (r0 + c) * GNT;
                const bool ok = x1 < g.H && c0 < g.W;
                const VT uv = gather_load<VT>(gu, gather_off<T>(gu, ok, x1 * g.W + c0));
                va[c] += uv;
                vb[c] += dn * uv;
              }
            }
          }
        }
#pragma unroll
        for (int c = 0; c < CR; ++c) {
          const int x1 = tid + (r0 + c) * GNT;
          if (x1 < N) {
#pragma unroll
            for (int v = 0; v < V; ++v) buf[v * VS + x1] = cplx<T>{va[c][v], vb[c][v]};
          }
        }
      }
    }
  }
  for (int u0 = 0; !row_vec && u0 < EPT; u0 += CHK) {
    cplx<T> val[CHK];
#pragma unroll
    for (int c = 0; c < CHK; ++c) val[c] = cplx<T>{T(0), T(0)};
    if constexpr (MODE == 0) {
      // Zero-reading gathers (the alpha field reads zero when absent): the
      // chunk's loads of all four arrays are in flight together.
#pragma unroll
      for (int c = 0; c < CHK; ++c) {
        const int e = tid + (u0 + c) * GNT;
        const int x1 = e / V, v = e - (e / V) * V;
        const int x2 = c0 + v;
        const bool ok = (V * N % GNT == 0 || e < V * N) && (N % V == 0 || x2 < N);
        const int o = x1 * N + x2;
        const int orf = gather_off<T>(gar, ok, o), oal = gather_off<T>(gaa, ok && g.has_alpha, o);
        val[c].re = gather_load<T>(gar, orf) * gather_load<T>(gxr, orf);
        val[c].im = gather_load<T>(gaa, oal) * gather_load<T>(gxa, oal);
      }
    } else if (g.ucomplex) {
      for (int pl = 0; pl < g.uplanes; ++pl) {
        const cplx<T>* up = reinterpret_cast<const cplx<T>*>(u) + pl * plane;
        const Gather gu = make_gather(up, plane);
#pragma unroll
        for (int c = 0; c < CHK; ++c) {
          const int e = tid + (u0 + c) * GNT;
          const int x1 = e / V, v = e - (e / V) * V;
          const int x2 = c0 + v;
          const bool ok = (V * N % GNT == 0 || e < V * N) && x1 < g.H && x2 < g.W;
          val[c] = cadd(val[c], gather_load<cplx<T>>(gu, gather_off<cplx<T>>(gu, ok, x1 * g.W + x2)));
        }
      }
    } else {
      for (int pl = 0; pl < g.uplanes; ++pl) {
        const Gather gu = make_gather(u + pl * plane, plane);
        const T dn = g.has_alpha ? dnu[pl % g.E] : T(0);
#pragma unroll
        for (int c = 0; c < CHK; ++c) {
          const int e = tid + (u0 + c) * GNT;
          const int x1 = e / V, v = e - (e / V) * V;
          const int x2 = c0 + v;
          const bool ok = (V * N % GNT == 0 || e < V * N) && x1 < g.H && x2 < g.W;
          const T uv = gather_load<T>(gu, gather_off<T>(gu, ok, x1 * g.W + x2));
          val[c].re += uv;
          val[c].im += dn * uv;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CHK; ++c) {
      const int e = tid + (u0 + c) * GNT;
      if (e >= V * N) continue;
      const int x1 = e / V, v = e - (e / V) * V;
      buf[v * VS + x1] = val[c];
    }
  }
#pragma unroll
  for (int u2 = 0; u2 < TPT; ++u2) {
    const int e = tid + u2 * GNT;
    if (e < N) tw[e] = twr[u2];
  }
  __syncthreads();
  fft_lds<T, N, V, VS, GNT, false, kFewVectorPlan<N, V, GNT>>(buf, tw, tid);
  for (int e = tid; e < V * N; e += GNT) {
    const int x1 = e / V, v = e - (e / V) * V;
    if (N % V == 0 || c0 + v < N) Tw[(int64_t)x1 * N + c0 + v] = buf[v * VS + x1];
  }
}

// grid (N / 2 + 1, S).  MODE 0: sky = exp(log s) cropped to (E, H, W);
// MODE 1: g_ref = amp_ref HT(u_ref), g_alpha = amp_alpha HT(u_alpha) on N x N;
// MODE 2: tangent ds_e = sky_e (HT(a) + HT(b) dnu_e) on (E, H, W);
// MODE 3: MODE 1 plus the amplitude-field and offset gradients (GpHyper):
// g_amp_ref = xi_ref HT(u_ref), g_amp_alpha = xi_alpha HT(u_alpha), and
// (d zm, d mu) = (HT(u_ref), HT(u_alpha)) at k = 0, i.e. the sums of u.
// Sample blockIdx.y reads work[s] and writes output s; the tangent's sky is
// the same base-point sky for every sample when g.sky_shared.
template <typename T, int N, int MODE>
__global__ __launch_bounds__(GNR) void gp_rows(GpGeom g, const cplx<T>* __restrict__ Tw,
                                               const cplx<T>* __restrict__ twg,
                                               const T* __restrict__ dnu,
                                               const T* __restrict__ amp_ref,
                                               const T* __restrict__ amp_alpha,
                                               const T* __restrict__ sky_in,
                                               T* __restrict__ out_a, T* __restrict__ out_b,
                                               GpHyper hy) {
  constexpr bool ADJ = MODE == 1 || MODE == 3;
  constexpr int VS = lds_stride(N, 2);
  __shared__ cplx<T> buf[2 * VS];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x;
  const int r1 = (N - r0) % N;
  const int64_t smp = blockIdx.y;
  if (!(MODE == 2 && g.work_shared)) Tw += smp * N * N;
  const T* xi_a = (const T*)hy.xi_ref;
  const T* xi_b = (const T*)hy.xi_alpha;
  T* ga_a = (T*)hy.g_amp_ref;
  T* ga_b = (T*)hy.g_amp_alpha;
  T* goff = (T*)hy.g_off;
  if constexpr (ADJ) {
    out_a += smp * N * N;
    if (g.has_alpha) out_b += smp * N * N;
    if constexpr (MODE == 3) {
      xi_a += smp * N * N;
      ga_a += smp * N * N;
      goff += smp * 2;
      if (g.has_alpha) {
        xi_b += smp * N * N;
        ga_b += smp * N * N;
      }
    }
  } else {
    out_a += smp * g.E * g.H * g.W;
    if (MODE == 2 && !g.sky_shared) sky_in += smp * g.E * g.H * g.W;
  }
  constexpr int TPT = (N + GNR - 1) / GNR;
  constexpr int EPT = (2 * N + GNR - 1) / GNR;
  cplx<T> twr[TPT], rv[EPT];
#pragma unroll
  for (int u2 = 0; u2 < EPT; ++u2) {
    const int e = tid + u2 * GNR;
    if (e < 2 * N) {
      const int v = e / N, k = e - (e / N) * N;
      rv[u2] = Tw[(int64_t)(v ? r1 : r0) * N + k];
    }
  }
#pragma unroll
  for (int u2 = 0; u2 < TPT; ++u2) {
    const int e = tid + u2 * GNR;
    if (e < N) twr[u2] = twg[e];
  }
  // MODE 1: the amplitudes of the epilogue are fetched with the rows (zero-
  // reading gathers; alpha reads zero when absent), so their latency
  // overlaps the transform.
  T ampa[ADJ ? EPT : 1], ampb[ADJ ? EPT : 1];
  T xia[MODE == 3 ? EPT : 1], xib[MODE == 3 ? EPT : 1];
  if constexpr (ADJ) {
    const Gather ga = make_gather(amp_ref, (int64_t)N * N);
    const Gather gb = make_gather(amp_alpha, g.has_alpha ? (int64_t)N * N : 0);
    const Gather gxa = make_gather(xi_a, MODE == 3 ? (int64_t)N * N : 0);
    const Gather gxb = make_gather(xi_b, MODE == 3 && g.has_alpha ? (int64_t)N * N : 0);
#pragma unroll
    for (int u2 = 0; u2 < EPT; ++u2) {
      const int e = tid + u2 * GNR;
      const int v = e / N, k = e - (e / N) * N;
      const bool ok = 2 * N % GNR == 0 || e < 2 * N;
      const int o = (v ? r1 : r0) * N + k;
      ampa[u2] = gather_load<T>(ga, gather_off<T>(ga, ok, o));
      ampb[u2] = gather_load<T>(gb, gather_off<T>(gb, ok && g.has_alpha, o));
      if constexpr (MODE == 3) {
        xia[u2] = gather_load<T>(gxa, gather_off<T>(gxa, ok, o));
        xib[u2] = gather_load<T>(gxb, gather_off<T>(gxb, ok && g.has_alpha, o));
      }
    }
  }
#pragma unroll
  for (int u2 = 0; u2 < EPT; ++u2) {
    const int e = tid + u2 * GNR;
    if (e < 2 * N) {
      const int v = e / N, k = e - (e / N) * N;
      buf[v * VS + k] = rv[u2];
    }
  }
#pragma unroll
  for (int u2 = 0; u2 < TPT; ++u2) {
    const int e = tid + u2 * GNR;
    if (e < N) tw[e] = twr[u2];
  }
  __syncthreads();
  fft_lds<T, N, 2, VS, GNR, false, kFewVectorPlan<N, 2, GNR>>(buf, tw, tid);
  const int nrows = (r0 == r1) ? 1 : 2;
#pragma unroll
  for (int u2 = 0; u2 < EPT; ++u2) {
    const int e = tid + u2 * GNR;
    if (e >= nrows * N) continue;
    const int v = e / N, k = e - (e / N) * N;
    const int x = v ? r1 : r0;
    const cplx<T> z = buf[v * VS + k];
    const cplx<T> zp = buf[(1 - v) * VS + (k == 0 ? 0 : N - k)];
    const T hta = T(0.5) * (z.re + zp.re + z.im - zp.im);
    const T htb = T(0.5) * (z.im + zp.im - z.re + zp.re);
    if constexpr (ADJ) {
      const int64_t o = (int64_t)x * N + k;
      out_a[o] = ampa[u2] * hta;
      if (g.has_alpha) out_b[o] = ampb[u2] * htb;
      if constexpr (MODE == 3) {
        ga_a[o] = xia[u2] * hta;
        if (g.has_alpha) ga_b[o] = xib[u2] * htb;
        if (x == 0 && k == 0) {
          goff[0] = hta;
          goff[1] = g.has_alpha ? htb : T(0);
        }
      }
    } else {
      if (x >= g.H || k >= g.W) continue;
      for (int en = 0; en < g.E; ++en) {
        const int64_t o = ((int64_t)en * g.H + x) * g.W + k;
        if constexpr (MODE == 0) {
          T ls = T(g.zm) + hta;
          if (g.has_alpha) ls += (htb + T(g.mu)) * dnu[en];
          out_a[o] = exp(ls);
        } else {
          T dl = hta;
          if (g.has_alpha) dl += htb * dnu[en];
          out_a[o] = sky_in[o] * dl;
        }
      }
    }
  }
}

// u_e(H, W) = sky_e * cot_e: the cotangent of log s per energy, the planes
// the adjoint column pass folds over the energies (the transpose of the
// energy broadcast).
template <typename T>
__global__ __launch_bounds__(GNT) void gp_fold(GpGeom g, const T* __restrict__ sky,
                                               const T* __restrict__ cot, T* __restrict__ u) {
  const int64_t n = (int64_t)g.E * g.H * g.W;
  for (int64_t i = blockIdx.x * (int64_t)GNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * GNT)
    u[i] = sky[i] * cot[i];
}

// ====================================================================== host
// GP grid sizes: powers of two, 768, and the 7-smooth grids 567 and 1134 that
// ducc0's good_size gives for the demos' padded grids (int(512 * 1.1075),
// int(1024 * 1.1075); jubik/sky_models.py:306).
#define JUBIK_GP_LENGTHS(X) \
  X(16) X(32) X(64) X(128) X(256) X(512) X(567) X(768) X(1024) X(1134) X(2048)

bool gp_length_supported(int N) {
  switch (N) {
#define JUBIK_CASE(l) case l:
    JUBIK_GP_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    return true;
    default:
      return false;
  }
}

template <typename T, int N>
struct GpOps {
  static constexpr int V = GpTile<T, N>::V;
  using C = cplx<T>;
  static int forward(const GpGeom& g, const GpBuffers& b, const void* amp_ref,
                     const void* amp_alpha, const void* xi_ref, const void* xi_alpha,
                     const void* dnu, void* sky, hipStream_t s) {
    {
      TimedLaunch _t("gp_cols_fwd", s);
      gp_cols<T, N, 0><<<dim3((N + V - 1) / V, g.S), GNT, 0, s>>>(g, (const T*)amp_ref, (const T*)amp_alpha,
                                             (const T*)xi_ref, (const T*)xi_alpha, nullptr,
                                             nullptr, (const C*)b.tw, (C*)b.work);
    }
    {
      TimedLaunch _t("gp_rows_fwd", s);
      gp_rows<T, N, 0><<<dim3(N / 2 + 1, g.S), GNR, 0, s>>>(g, (const C*)b.work, (const C*)b.tw,
                                                 (const T*)dnu, nullptr, nullptr, nullptr,
                                                 (T*)sky, nullptr, GpHyper{});
    }
    return check_hip(hipGetLastError(), "gp_forward");
  }
  static int tangent(const GpGeom& g, const GpBuffers& b, const void* amp_ref,
                     const void* amp_alpha, const void* v_ref, const void* v_alpha,
                     const void* dnu, const void* sky, void* dsky, hipStream_t s) {
    {
      TimedLaunch _t("gp_cols_fwd", s);
      gp_cols<T, N, 0><<<dim3((N + V - 1) / V, g.work_shared ? 1 : g.S), GNT, 0, s>>>(
          g, (const T*)amp_ref, (const T*)amp_alpha, (const T*)v_ref, (const T*)v_alpha, nullptr,
          nullptr, (const C*)b.tw, (C*)b.work);
    }
    {
      TimedLaunch _t("gp_rows_tangent", s);
      gp_rows<T, N, 2><<<dim3(N / 2 + 1, g.S), GNR, 0, s>>>(g, (const C*)b.work, (const C*)b.tw,
                                                 (const T*)dnu, nullptr, nullptr, (const T*)sky,
                                                 (T*)dsky, nullptr, GpHyper{});
    }
    return check_hip(hipGetLastError(), "gp_tangent");
  }
  static int adjoint_u(const GpGeom& g, const GpBuffers& b, const void* amp_ref,
                       const void* amp_alpha, const void* dnu, const void* u, void* g_ref,
                       void* g_alpha, hipStream_t s, GpHyper hy) {
    {
      TimedLaunch _t("gp_cols_adj", s);
      constexpr int VA = GpTile<T, N, 1>::V;
      gp_cols<T, N, 1><<<dim3((N + VA - 1) / VA, g.S), GNT, 0, s>>>(
          g, nullptr, nullptr, nullptr, nullptr, (const T*)u, (const T*)dnu, (const C*)b.tw,
          (C*)b.work);
    }
    if (hy.g_amp_ref) {
      TimedLaunch _t("gp_rows_adj_hyper", s);
      gp_rows<T, N, 3><<<dim3(N / 2 + 1, g.S), GNR, 0, s>>>(g, (const C*)b.work, (const C*)b.tw, nullptr,
                                                 (const T*)amp_ref, (const T*)amp_alpha, nullptr,
                                                 (T*)g_ref, (T*)g_alpha, hy);
    } else {
      TimedLaunch _t("gp_rows_adj", s);
      gp_rows<T, N, 1><<<dim3(N / 2 + 1, g.S), GNR, 0, s>>>(g, (const C*)b.work, (const C*)b.tw, nullptr,
                                                 (const T*)amp_ref, (const T*)amp_alpha, nullptr,
                                                 (T*)g_ref, (T*)g_alpha, GpHyper{});
    }
    return check_hip(hipGetLastError(), "gp_adjoint");
  }
  static int fold(const GpGeom& g, const void* sky, const void* cot, void* u, hipStream_t s) {
    const int64_t n = (int64_t)g.E * g.H * g.W;
    const int blocks = (int)std::min<int64_t>((n + GNT - 1) / GNT, 4096);
    {
      TimedLaunch _t("gp_fold", s);
      gp_fold<T><<<blocks, GNT, 0, s>>>(g, (const T*)sky, (const T*)cot, (T*)u);
    }
    return check_hip(hipGetLastError(), "gp_fold");
  }
};

template <typename T, typename F>
static int gp_dispatch_len(int N, F&& f) {
  switch (N) {
#define JUBIK_CASE(l) \
  case l:             \
    return f(GpOps<T, l>{});
    JUBIK_GP_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    default:
      return fail(kUnsupported, "GP grid size " + std::to_string(N) + " is not compiled in");
  }
}

template <typename F>
static int gp_dispatch(int dtype, int N, F&& f) {
  if (dtype == 0) return gp_dispatch_len<float>(N, f);
  if (dtype == 1) return gp_dispatch_len<double>(N, f);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}

int gp_forward(int dtype, const GpGeom& g, const GpBuffers& b, const void* amp_ref,
               const void* amp_alpha, const void* xi_ref, const void* xi_alpha, const void* dnu,
               void* sky, hipStream_t s) {
  return gp_dispatch(dtype, g.N, [&](auto ops) {
    return decltype(ops)::forward(g, b, amp_ref, amp_alpha, xi_ref, xi_alpha, dnu, sky, s);
  });
}
int gp_tangent(int dtype, const GpGeom& g, const GpBuffers& b, const void* amp_ref,
               const void* amp_alpha, const void* v_ref, const void* v_alpha, const void* dnu,
               const void* sky, void* dsky, hipStream_t s) {
  return gp_dispatch(dtype, g.N, [&](auto ops) {
    return decltype(ops)::tangent(g, b, amp_ref, amp_alpha, v_ref, v_alpha, dnu, sky, dsky, s);
  });
}
int gp_adjoint_u(int dtype, const GpGeom& g, const GpBuffers& b, const void* amp_ref,
                 const void* amp_alpha, const void* dnu, const void* u, void* g_ref,
                 void* g_alpha, hipStream_t s, GpHyper hy) {
  if (g.has_alpha && !dnu) return fail(kInvalidArgument, "spectral-index inputs are NULL");
  return gp_dispatch(dtype, g.N, [&](auto ops) {
    return decltype(ops)::adjoint_u(g, b, amp_ref, amp_alpha, dnu, u, g_ref, g_alpha, s, hy);
  });
}
int gp_fold(int dtype, const GpGeom& g, const void* sky, const void* cot, void* u,
            hipStream_t s) {
  return gp_dispatch(dtype, g.N, [&](auto ops) { return decltype(ops)::fold(g, sky, cot, u, s); });
}

}  // namespace jubik
