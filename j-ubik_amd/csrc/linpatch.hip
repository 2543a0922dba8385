// Spatially variant PSF by linear patching (jubik/convolve.py:66-145) and
// its adjoint, as three fused passes per direction:
//
//   rows : gather two real window rows per complex vector (tent weights and
//          zero margins applied on the fly), FFT along the row, split the two
//          half spectra and store them column-major (R1[.][k][r]).
//   cols : FFT along the columns, multiply by the prepared kernel spectrum
//          (conjugated for the adjoint), inverse FFT, store row-major
//          (R2[.][f][k]).
//   out  : per output image row, rebuild the full spectra of two patches per
//          complex vector, inverse FFT, and sum the patch rows that cover the
//          row in a fixed order (gather-form overlap-add, no atomics), with a
//          fused epilogue (plain store, or exposure + Poisson residual + FP64
//          NLL partials, or the chain into the GP adjoint).
//
// Kernel spectra are prepared by the same rows/cols passes from the cut
// window, with the ifftshift applied as (-1)^(k1+k2) and the 1/(sum L^2)
// normalisation folded in (convolve.py:104-131 and 162-176; dvol cancels).
#include "linpatch.hpp"

namespace jubik {

enum RowMode { ROW_SIGNAL = 0, ROW_COTANGENT = 1, ROW_KERNEL_WIN = 2, ROW_KERNEL_FULL = 3 };
enum ColMode { COL_FWD = 0, COL_ADJ = 1, COL_KERNEL = 2 };

constexpr int NT = 256;
constexpr int MAXN = 64;

// ------------------------------------------------------------------ rows
// grid (ceil(nrows/2 / V), P, Z); Z = jobs (signal modes) or M*Ek (kernel).
template <typename T, int L, int MODE>
__global__ __launch_bounds__(NT) void lp_rows(LinGeom g, const T* __restrict__ src,
                                              const T* __restrict__ tent,
                                              const cplx<T>* __restrict__ twg,
                                              cplx<T>* __restrict__ R1, T scale, int nrows) {
  constexpr int V = Tile<T, L>::V;
  constexpr int L2 = L / 2 + 1;
  __shared__ cplx<T> buf[V * L];
  __shared__ cplx<T> tw[L];
  const int tid = threadIdx.x;
  const int p = blockIdx.y;
  const int z = blockIdx.z;
  const int q0 = blockIdx.x * V;
  const int ix = p % g.n, iy = p / g.n;
  const int sx = start_x(g, ix), sy = start_y(g, iy);

  int64_t r1base;
  const T* img;
  if constexpr (MODE == ROW_SIGNAL || MODE == ROW_COTANGENT) {
    const int job = z;
    r1base = (int64_t)job * g.P + p;
    int im = job;
    if (MODE == ROW_SIGNAL && g.x_shared) im = job % g.E;
    img = src + (int64_t)im * g.H * g.W;
  } else {
    const int mod = z / g.Ek, ek = z - mod * g.Ek;
    const int64_t spec = ((int64_t)mod * g.P + p) * g.Ek + ek;
    r1base = spec;
    img = src + spec * (MODE == ROW_KERNEL_WIN ? (int64_t)g.w * g.w : (int64_t)g.H * g.W);
  }

  load_twiddles(tw, twg, L, tid, NT);
  for (int e = tid; e < V * L; e += NT) {
    const int v = e / L, f = e - (e / L) * L;
    const int r0 = 2 * (q0 + v);
    T a = T(0), b = T(0);
    if (r0 < nrows) {
      if constexpr (MODE == ROW_SIGNAL) {
        const int c = f - g.m;
        if (c >= 0 && c < g.w) {
          const int j = sy + c - g.h;
          const int i0 = sx + r0 - g.h;
          if (j >= 0 && j < g.W) {
            const T wc = tent[c] * scale;
            if (i0 >= 0 && i0 < g.H) a = tent[r0] * wc * img[(int64_t)i0 * g.W + j];
            if (i0 + 1 >= 0 && i0 + 1 < g.H) b = tent[r0 + 1] * wc * img[(int64_t)(i0 + 1) * g.W + j];
          }
        }
      } else if constexpr (MODE == ROW_COTANGENT) {
        const int j = sy + f - g.h - g.m;
        const int i0 = sx + r0 - g.h - g.m;
        if (j >= 0 && j < g.W) {
          if (i0 >= 0 && i0 < g.H) a = img[(int64_t)i0 * g.W + j];
          if (i0 + 1 >= 0 && i0 + 1 < g.H) b = img[(int64_t)(i0 + 1) * g.W + j];
        }
      } else if constexpr (MODE == ROW_KERNEL_WIN) {
        const int c = f - g.m;
        if (c >= 0 && c < g.w) {
          a = img[(int64_t)r0 * g.w + c];
          b = img[(int64_t)(r0 + 1) * g.w + c];
        }
      } else {  // ROW_KERNEL_FULL: cut[r][c] = K[(r - d) mod H][(c - d) mod W]
        const int c = f - g.m;
        if (c >= 0 && c < g.w) {
          const int j = (c - g.d + g.W) % g.W;
          const int ia = (r0 - g.d + g.H) % g.H;
          const int ib = (r0 + 1 - g.d + g.H) % g.H;
          a = img[(int64_t)ia * g.W + j];
          b = img[(int64_t)ib * g.W + j];
        }
      }
    }
    buf[v * L + f] = cplx<T>{a, b};
  }
  __syncthreads();
  fft_lds<T, L, V, L, NT, false>(buf, tw, tid);

  cplx<T>* out = R1 + r1base * L2 * nrows;
  for (int e = tid; e < L2 * V; e += NT) {
    const int k = e / V, v = e - (e / V) * V;
    const int r0 = 2 * (q0 + v);
    if (r0 >= nrows) continue;
    const cplx<T> zk = buf[v * L + k];
    const cplx<T> zn = buf[v * L + (k == 0 ? 0 : L - k)];
    const cplx<T> A = {T(0.5) * (zk.re + zn.re), T(0.5) * (zk.im - zn.im)};
    const cplx<T> B = {T(0.5) * (zk.im + zn.im), T(0.5) * (zn.re - zk.re)};
    out[(int64_t)k * nrows + r0] = A;
    out[(int64_t)k * nrows + r0 + 1] = B;
  }
}

// ------------------------------------------------------------------ cols
// grid (ceil(L2 / V), P, Z).
template <typename T, int L, int MODE>
__global__ __launch_bounds__(NT) void lp_cols(LinGeom g, const cplx<T>* __restrict__ R1,
                                              const cplx<T>* __restrict__ KT,
                                              const cplx<T>* __restrict__ twg,
                                              cplx<T>* __restrict__ out) {
  constexpr int V = Tile<T, L>::V;
  constexpr int L2 = L / 2 + 1;
  __shared__ cplx<T> buf[V * L];
  __shared__ cplx<T> tw[L];
  __shared__ double red[NT / 64];
  __shared__ double total_s;
  const int tid = threadIdx.x;
  const int p = blockIdx.y;
  const int z = blockIdx.z;
  const int k0 = blockIdx.x * V;

  int64_t idx, spec;
  if constexpr (MODE == COL_KERNEL) {
    const int mod = z / g.Ek, ek = z - mod * g.Ek;
    spec = ((int64_t)mod * g.P + p) * g.Ek + ek;
    idx = spec;
  } else {
    idx = (int64_t)z * g.P + p;
    spec = spec_index(g, z, p);
  }
  const int nin = (MODE == COL_ADJ) ? L : g.w;
  const cplx<T>* in = R1 + idx * L2 * nin;

  load_twiddles(tw, twg, L, tid, NT);
  for (int e = tid; e < V * L; e += NT) {
    const int v = e / L, f = e - (e / L) * L;
    const int k = k0 + v;
    cplx<T> val = {T(0), T(0)};
    if (k < L2) {
      if constexpr (MODE == COL_ADJ) {
        val = in[(int64_t)k * nin + f];
      } else {
        const int rr = f - g.m;
        if (rr >= 0 && rr < g.w) val = in[(int64_t)k * nin + rr];
      }
    }
    buf[v * L + f] = val;
  }
  T factor = T(1);
  if constexpr (MODE == COL_KERNEL) {
    double s = 0.0;
    for (int rr = tid; rr < g.w; rr += NT) s += (double)in[rr].re;
    s = block_sum<NT>(s, red);
    if (tid == 0) total_s = s;
  }
  __syncthreads();
  if constexpr (MODE == COL_KERNEL) factor = T(1.0 / (total_s * (double)L * (double)L));
  fft_lds<T, L, V, L, NT, false>(buf, tw, tid);

  if constexpr (MODE == COL_KERNEL) {
    cplx<T>* kt = out + spec * L2 * L;
    for (int e = tid; e < V * L; e += NT) {
      const int v = e / L, f = e - (e / L) * L;
      const int k = k0 + v;
      if (k >= L2) continue;
      const T sgn = ((k + f) & 1) ? -factor : factor;
      kt[(int64_t)k * L + f] = cscale(buf[v * L + f], sgn);
    }
    return;
  } else {
    const cplx<T>* ks = KT + spec * L2 * L;
    for (int e = tid; e < V * L; e += NT) {
      const int v = e / L, f = e - (e / L) * L;
      const int k = k0 + v;
      if (k < L2) {
        const cplx<T> kv = ks[(int64_t)k * L + f];
        buf[v * L + f] = (MODE == COL_ADJ) ? cmulc(buf[v * L + f], kv) : cmul(buf[v * L + f], kv);
      }
    }
    __syncthreads();
    fft_lds<T, L, V, L, NT, true>(buf, tw, tid);
    const int nout = (MODE == COL_ADJ) ? g.w : L;
    const int foff = (MODE == COL_ADJ) ? g.m : 0;
    cplx<T>* o = out + idx * L2 * nout;
    for (int e = tid; e < nout * V; e += NT) {
      const int f = e / V, v = e - (e / V) * V;
      const int k = k0 + v;
      if (k < L2) o[(int64_t)f * L2 + k] = buf[v * L + f + foff];
    }
  }
}

// Rebuild z = A + iB from two half spectra (rows of R2) and store in LDS.
template <typename T, int L>
__device__ __forceinline__ cplx<T> hermitian_pair(const cplx<T>* a, const cplx<T>* b, int k) {
  const bool cj = k > L / 2;
  const int kk = cj ? L - k : k;
  cplx<T> A = a[kk];
  cplx<T> B = b ? b[kk] : cplx<T>{T(0), T(0)};
  if (cj) {
    A.im = -A.im;
    B.im = -B.im;
  }
  return {A.re - B.im, A.im + B.re};
}

// ------------------------------------------------------------------ fwd out
// Output columns are owned by threads (j = tid + t * NT), accumulators live
// in registers, so every pixel's sum runs in a fixed order.
constexpr int JPT = 16;  // W <= JPT * NT

// grid (H, J).
template <typename T, int L, int EPI>
__global__ __launch_bounds__(NT) void lp_fwd_out(LinGeom g, const cplx<T>* __restrict__ R2,
                                                 const cplx<T>* __restrict__ twg,
                                                 T* __restrict__ y, const T* __restrict__ expo,
                                                 const int* __restrict__ counts,
                                                 T* __restrict__ resid,
                                                 double* __restrict__ partial) {
  constexpr int V = Tile<T, L>::V;
  constexpr int L2 = L / 2 + 1;
  __shared__ cplx<T> buf[V * L];
  __shared__ cplx<T> tw[L];
  __shared__ int lst_ix[MAXN], lst_f[MAXN];
  __shared__ int nc_s;
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int job = blockIdx.y;
  const int R = i + g.h + g.m;

  load_twiddles(tw, twg, L, tid, NT);
  if (tid == 0) {
    int nc = 0;
    for (int ix = 0; ix < g.n; ++ix) {
      const int s = start_x(g, ix);
      if (R >= s && R < s + L) {
        lst_ix[nc] = ix;
        lst_f[nc] = R - s;
        ++nc;
      }
    }
    nc_s = nc;
  }
  T acc[JPT];
#pragma unroll
  for (int t = 0; t < JPT; ++t) acc[t] = T(0);
  __syncthreads();
  const int npair = (g.n + 1) / 2;
  const int nvec = nc_s * npair;
  const cplx<T>* base = R2 + (int64_t)job * g.P * L * L2;

  for (int v0 = 0; v0 < nvec; v0 += V) {
    const int cnt = min(V, nvec - v0);
    for (int e = tid; e < cnt * L; e += NT) {
      const int v = e / L, k = e - (e / L) * L;
      const int vv = v0 + v;
      const int ci = vv / npair, q = vv - (vv / npair) * npair;
      const int ix = lst_ix[ci], f = lst_f[ci];
      const int iya = 2 * q, iyb = 2 * q + 1;
      const cplx<T>* ra = base + ((int64_t)(iya * g.n + ix) * L + f) * L2;
      const cplx<T>* rb = iyb < g.n ? base + ((int64_t)(iyb * g.n + ix) * L + f) * L2 : nullptr;
      buf[v * L + k] = hermitian_pair<T, L>(ra, rb, k);
    }
    __syncthreads();
    fft_lds<T, L, V, L, NT, true>(buf, tw, tid);
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NT;
      if (j < g.W) {
        const int C = j + g.h + g.m;
        T s = T(0);
        for (int v = 0; v < cnt; ++v) {
          const int vv = v0 + v;
          const int q = vv % npair;
          const int ca = C - start_y(g, 2 * q);
          if ((unsigned)ca < (unsigned)L) s += buf[v * L + ca].re;
          if (2 * q + 1 < g.n) {
            const int cb = C - start_y(g, 2 * q + 1);
            if ((unsigned)cb < (unsigned)L) s += buf[v * L + cb].im;
          }
        }
        acc[t] += s;
      }
    }
    __syncthreads();
  }

  const int64_t row = ((int64_t)job * g.H + i) * g.W;
  if constexpr (EPI == 0) {
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NT;
      if (j < g.W) y[row + j] = acc[t];
    }
  } else if constexpr (EPI == 2) {
    // Metric: cotangent expo * (d lam / lam) on read-out pixels, y = lam.
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NT;
      if (j < g.W) {
        const T ex = expo[row + j];
        resid[row + j] = counts[row + j] >= 0 ? ex * (ex * acc[t]) / y[row + j] : T(0);
      }
    }
  } else {
    double nll = 0.0;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NT;
      if (j < g.W) {
        const T ex = expo[row + j];
        const T lam = ex * acc[t];
        const int d = counts[row + j];
        T r = T(0);
        if (d >= 0) {
          const double dl = (double)lam;
          nll += dl - (double)d * log(dl);
          r = ex * (T(1) - T(d) / lam);
        }
        resid[row + j] = r;
        if (y) y[row + j] = lam;
      }
    }
    nll = block_sum<NT>(nll, red);
    if (tid == 0) partial[(int64_t)job * g.H + i] = nll;
  }
}

// ------------------------------------------------------------------ adj out
// EPI 0: grid (H, nimg); gx[img] = scale * sum over the jobs of img.
// EPI 1: grid (H, 1);   u = sum_jobs scale * sky[e] * xbar_job * (1, dnu[e]),
//        stored as complex (u_ref, u_alpha) rows of width W.
template <typename T, int L, int EPI>
__global__ __launch_bounds__(NT) void lp_adj_out(LinGeom g, const cplx<T>* __restrict__ R2,
                                                 const cplx<T>* __restrict__ twg,
                                                 const T* __restrict__ tent,
                                                 T* __restrict__ gx, T scale,
                                                 const T* __restrict__ sky,
                                                 const T* __restrict__ dnu,
                                                 cplx<T>* __restrict__ u) {
  constexpr int V = Tile<T, L>::V;
  constexpr int L2 = L / 2 + 1;
  __shared__ cplx<T> buf[V * L];
  __shared__ cplx<T> tw[L];
  __shared__ int lst_ix[MAXN], lst_r[MAXN];
  __shared__ int nc_s;
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int Rp = i + g.h;

  load_twiddles(tw, twg, L, tid, NT);
  if (tid == 0) {
    int nc = 0;
    for (int ix = 0; ix < g.n; ++ix) {
      const int s = start_x(g, ix);
      if (Rp >= s && Rp < s + g.w) {
        lst_ix[nc] = ix;
        lst_r[nc] = Rp - s;
        ++nc;
      }
    }
    nc_s = nc;
  }
  T acc[JPT], acc2[JPT], tmp[JPT];
#pragma unroll
  for (int t = 0; t < JPT; ++t) acc[t] = acc2[t] = T(0);
  __syncthreads();
  const int npair = (g.n + 1) / 2;
  const int nvec = nc_s * npair;

  int jb = 0, je = g.J, jstep = 1;
  if (EPI == 0) {
    jb = blockIdx.y;
    if (g.x_shared) {
      jstep = g.E;
    } else {
      je = blockIdx.y + 1;
    }
  }

  for (int job = jb; job < je; job += jstep) {
    const cplx<T>* base = R2 + (int64_t)job * g.P * g.w * L2;
#pragma unroll
    for (int t = 0; t < JPT; ++t) tmp[t] = T(0);
    for (int v0 = 0; v0 < nvec; v0 += V) {
      const int cnt = min(V, nvec - v0);
      for (int e = tid; e < cnt * L; e += NT) {
        const int v = e / L, k = e - (e / L) * L;
        const int vv = v0 + v;
        const int ci = vv / npair, q = vv - (vv / npair) * npair;
        const int ix = lst_ix[ci], r = lst_r[ci];
        const int iya = 2 * q, iyb = 2 * q + 1;
        const cplx<T>* ra = base + ((int64_t)(iya * g.n + ix) * g.w + r) * L2;
        const cplx<T>* rb = iyb < g.n ? base + ((int64_t)(iyb * g.n + ix) * g.w + r) * L2 : nullptr;
        buf[v * L + k] = hermitian_pair<T, L>(ra, rb, k);
      }
      __syncthreads();
      fft_lds<T, L, V, L, NT, true>(buf, tw, tid);
#pragma unroll
      for (int t = 0; t < JPT; ++t) {
        const int j = tid + t * NT;
        if (j < g.W) {
          const int Cp = j + g.h;
          T s = T(0);
          for (int v = 0; v < cnt; ++v) {
            const int vv = v0 + v;
            const int ci = vv / npair, q = vv - (vv / npair) * npair;
            const T wr = tent[lst_r[ci]];
            const int ca = Cp - start_y(g, 2 * q);
            if ((unsigned)ca < (unsigned)g.w) s += wr * tent[ca] * buf[v * L + g.m + ca].re;
            if (2 * q + 1 < g.n) {
              const int cb = Cp - start_y(g, 2 * q + 1);
              if ((unsigned)cb < (unsigned)g.w) s += wr * tent[cb] * buf[v * L + g.m + cb].im;
            }
          }
          tmp[t] += s;
        }
      }
      __syncthreads();
    }
    if (EPI == 0) {
#pragma unroll
      for (int t = 0; t < JPT; ++t) acc[t] += tmp[t];
    } else {
      const int e = job % g.E;
      const T dn = dnu ? dnu[e] : T(0);
      const T* srow = sky + ((int64_t)e * g.H + i) * g.W;
#pragma unroll
      for (int t = 0; t < JPT; ++t) {
        const int j = tid + t * NT;
        if (j < g.W) {
          const T val = scale * srow[j] * tmp[t];
          acc[t] += val;
          acc2[t] += dn * val;
        }
      }
    }
  }
  if (EPI == 0) {
    const int64_t row = ((int64_t)blockIdx.y * g.H + i) * g.W;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NT;
      if (j < g.W) gx[row + j] = scale * acc[t];
    }
  } else {
    const int64_t row = (int64_t)i * g.W;
#pragma unroll
    for (int t = 0; t < JPT; ++t) {
      const int j = tid + t * NT;
      if (j < g.W) u[row + j] = cplx<T>{acc[t], acc2[t]};
    }
  }
}

}  // namespace jubik

// ====================================================================== host
namespace jubik {

#define JUBIK_LP_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(224) X(256) X(324) X(384)

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

static int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return check_hip(e, what);
  return kOk;
}

template <typename T, int L>
struct LinOps {
  static constexpr int V = Tile<T, L>::V;
  static constexpr int L2 = L / 2 + 1;
  using C = cplx<T>;

  static int prepare(const LinGeom& g, const LinBuffers& b, const void* kernel, int layout,
                     hipStream_t s) {
    dim3 gr((g.w / 2 + V - 1) / V, g.P, g.M * g.Ek);
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
    dim3 gc((L2 + V - 1) / V, g.P, g.M * g.Ek);
    {
      TimedLaunch _t("lp_cols_kernel", s);
      lp_cols<T, L, COL_KERNEL><<<gc, NT, 0, s>>>(g, (const C*)b.R1, nullptr, (const C*)b.tw,
                                                  (C*)b.KT);
    }
    return launch_status("linpatch_prepare");
  }

  static void fwd_front(const LinGeom& g, const LinBuffers& b, const void* x, double scale,
                        hipStream_t s) {
    dim3 gr((g.w / 2 + V - 1) / V, g.P, g.J);
    {
      TimedLaunch _t("lp_rows_fwd", s);
      lp_rows<T, L, ROW_SIGNAL><<<gr, NT, 0, s>>>(g, (const T*)x, (const T*)b.tent, (const C*)b.tw,
                                                  (C*)b.R1, T(scale), g.w);
    }
    dim3 gc((L2 + V - 1) / V, g.P, g.J);
    {
      TimedLaunch _t("lp_cols_fwd", s);
      lp_cols<T, L, COL_FWD><<<gc, NT, 0, s>>>(g, (const C*)b.R1, (const C*)b.KT, (const C*)b.tw,
                                               (C*)b.R2);
    }
  }

  static void adj_front(const LinGeom& g, const LinBuffers& b, const void* gy, hipStream_t s) {
    dim3 gr((L / 2 + V - 1) / V, g.P, g.J);
    {
      TimedLaunch _t("lp_rows_adj", s);
      lp_rows<T, L, ROW_COTANGENT><<<gr, NT, 0, s>>>(g, (const T*)gy, (const T*)b.tent,
                                                     (const C*)b.tw, (C*)b.R1, T(1), L);
    }
    dim3 gc((L2 + V - 1) / V, g.P, g.J);
    {
      TimedLaunch _t("lp_cols_adj", s);
      lp_cols<T, L, COL_ADJ><<<gc, NT, 0, s>>>(g, (const C*)b.R1, (const C*)b.KT, (const C*)b.tw,
                                               (C*)b.R2);
    }
  }

  static int forward(const LinGeom& g, const LinBuffers& b, const void* x, void* y, double scale,
                     hipStream_t s) {
    fwd_front(g, b, x, scale, s);
    {
      TimedLaunch _t("lp_fwd_out", s);
      lp_fwd_out<T, L, 0><<<dim3(g.H, g.J), NT, 0, s>>>(g, (const C*)b.R2, (const C*)b.tw, (T*)y,
                                                        nullptr, nullptr, nullptr, nullptr);
    }
    return launch_status("linpatch_forward");
  }

  static int forward_poisson(const LinGeom& g, const LinBuffers& b, const void* x, double scale,
                             const void* expo, const int* counts, void* resid, void* lam,
                             double* partial, hipStream_t s) {
    fwd_front(g, b, x, scale, s);
    {
      TimedLaunch _t("lp_fwd_out_poisson", s);
      lp_fwd_out<T, L, 1><<<dim3(g.H, g.J), NT, 0, s>>>(g, (const C*)b.R2, (const C*)b.tw, (T*)lam,
                                                        (const T*)expo, counts, (T*)resid, partial);
    }
    return launch_status("linpatch_forward_poisson");
  }

  static int forward_metric(const LinGeom& g, const LinBuffers& b, const void* dx, double scale,
                            const void* expo, const int* counts, const void* lam, void* cot,
                            hipStream_t s) {
    fwd_front(g, b, dx, scale, s);
    {
      TimedLaunch _t("lp_fwd_out_metric", s);
      lp_fwd_out<T, L, 2><<<dim3(g.H, g.J), NT, 0, s>>>(g, (const C*)b.R2, (const C*)b.tw,
                                                        (T*)lam, (const T*)expo, counts, (T*)cot,
                                                        nullptr);
    }
    return launch_status("linpatch_forward_metric");
  }

  static int adjoint(const LinGeom& g, const LinBuffers& b, const void* gy, void* gx,
                     double scale, hipStream_t s) {
    adj_front(g, b, gy, s);
    const int nimg = g.x_shared ? g.E : g.J;
    {
      TimedLaunch _t("lp_adj_out", s);
      lp_adj_out<T, L, 0><<<dim3(g.H, nimg), NT, 0, s>>>(g, (const C*)b.R2, (const C*)b.tw,
                                                         (const T*)b.tent, (T*)gx, T(scale),
                                                         nullptr, nullptr, nullptr);
    }
    return launch_status("linpatch_adjoint");
  }

  static int adjoint_gp(const LinGeom& g, const LinBuffers& b, const void* gy, double scale,
                        const void* sky, const void* dnu, void* u, hipStream_t s) {
    adj_front(g, b, gy, s);
    {
      TimedLaunch _t("lp_adj_out_gp", s);
      lp_adj_out<T, L, 1><<<dim3(g.H, 1), NT, 0, s>>>(g, (const C*)b.R2, (const C*)b.tw,
                                                      (const T*)b.tent, nullptr, T(scale),
                                                      (const T*)sky, (const T*)dnu, (C*)u);
    }
    return launch_status("linpatch_adjoint_gp");
  }
};

template <template <typename, int> class Ops, typename T, typename F>
static int dispatch_len(int L, F&& f) {
  switch (L) {
#define JUBIK_CASE(l) \
  case l:             \
    return f(Ops<T, l>{});
    JUBIK_LP_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    default:
      return fail(kUnsupported, "FFT length " + std::to_string(L) + " is not compiled in");
  }
}

template <typename F>
static int dispatch(int dtype, int L, F&& f) {
  if (dtype == 0) return dispatch_len<LinOps, float>(L, f);
  if (dtype == 1) return dispatch_len<LinOps, double>(L, f);
  return fail(kInvalidArgument, "dtype must be 0 (float32) or 1 (float64)");
}

int linpatch_prepare(int dtype, const LinGeom& g, const LinBuffers& b, const void* kernel,
                     int layout, hipStream_t s) {
  return dispatch(dtype, g.L, [&](auto ops) { return decltype(ops)::prepare(g, b, kernel, layout, s); });
}
int linpatch_forward(int dtype, const LinGeom& g, const LinBuffers& b, const void* x, void* y,
                     double scale, hipStream_t s) {
  return dispatch(dtype, g.L, [&](auto ops) { return decltype(ops)::forward(g, b, x, y, scale, s); });
}
int linpatch_forward_poisson(int dtype, const LinGeom& g, const LinBuffers& b, const void* x,
                             double scale, const void* expo, const int* counts, void* resid,
                             void* lam, double* partial, hipStream_t s) {
  return dispatch(dtype, g.L, [&](auto ops) {
    return decltype(ops)::forward_poisson(g, b, x, scale, expo, counts, resid, lam, partial, s);
  });
}
int linpatch_forward_metric(int dtype, const LinGeom& g, const LinBuffers& b, const void* dx,
                            double scale, const void* expo, const int* counts, const void* lam,
                            void* cot, hipStream_t s) {
  return dispatch(dtype, g.L, [&](auto ops) {
    return decltype(ops)::forward_metric(g, b, dx, scale, expo, counts, lam, cot, s);
  });
}
int linpatch_adjoint(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy, void* gx,
                     double scale, hipStream_t s) {
  return dispatch(dtype, g.L, [&](auto ops) { return decltype(ops)::adjoint(g, b, gy, gx, scale, s); });
}
int linpatch_adjoint_gp(int dtype, const LinGeom& g, const LinBuffers& b, const void* gy,
                        double scale, const void* sky, const void* dnu, void* u, hipStream_t s) {
  return dispatch(dtype, g.L, [&](auto ops) {
    return decltype(ops)::adjoint_gp(g, b, gy, scale, sky, dnu, u, s);
  });
}

}  // namespace jubik
