// Stand-alone primitives of jubik/convolve.py: slice_patches (28-63) with its
// transpose, and the FFT convolution convolve (148-176) of real square
// arrays over the last two axes.  The fused PSF operator does not use them;
// they make the public functions available on the GPU one by one.
#include "primitives.hpp"

namespace jubik {

constexpr int PNT = 256;

// ------------------------------------------------------------ slicing
struct SliceGeom {
  int B, H, W, n;
  int dx, dy, hx, hy, wx, wy;  // strides, zero pad, window sizes
  int limx, limy;              // clamps of the window starts (dynamic_slice)
};

__device__ __forceinline__ int sl_start(int i, int d, int lim) { return min(i * d, lim); }

// out (P, B, wx, wy) = windows of the zero-padded x (B, H, W).
template <typename T>
__global__ __launch_bounds__(PNT) void k_slice(SliceGeom g, const T* __restrict__ x,
                                               T* __restrict__ out, int64_t total) {
  for (int64_t e = blockIdx.x * (int64_t)PNT + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * PNT) {
    const int c = (int)(e % g.wy);
    int64_t rest = e / g.wy;
    const int r = (int)(rest % g.wx);
    rest /= g.wx;
    const int b = (int)(rest % g.B);
    const int p = (int)(rest / g.B);
    const int i = sl_start(p % g.n, g.dx, g.limx) + r - g.hx;
    const int j = sl_start(p / g.n, g.dy, g.limy) + c - g.hy;
    out[e] = (i >= 0 && i < g.H && j >= 0 && j < g.W) ? x[((int64_t)b * g.H + i) * g.W + j]
                                                      : T(0);
  }
}

// x (B, H, W) = sum of the windows covering each pixel (transpose of k_slice),
// gathered per pixel in a fixed order.
template <typename T>
__global__ __launch_bounds__(PNT) void k_slice_adj(SliceGeom g, const T* __restrict__ patches,
                                                   T* __restrict__ x, int64_t total) {
  for (int64_t e = blockIdx.x * (int64_t)PNT + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * PNT) {
    const int j = (int)(e % g.W);
    const int64_t rest = e / g.W;
    const int i = (int)(rest % g.H);
    const int b = (int)(rest / g.H);
    const int ip = i + g.hx, jp = j + g.hy;
    T acc = T(0);
    for (int ix = ip >= g.wx ? (ip - g.wx) / g.dx + 1 : 0; ix < g.n; ++ix) {
      const int sx = sl_start(ix, g.dx, g.limx);
      if (sx > ip) break;
      if (ip >= sx + g.wx) continue;
      for (int iy = jp >= g.wy ? (jp - g.wy) / g.dy + 1 : 0; iy < g.n; ++iy) {
        const int sy = sl_start(iy, g.dy, g.limy);
        if (sy > jp) break;
        if (jp >= sy + g.wy) continue;
        const int p = iy * g.n + ix;
        acc += patches[(((int64_t)p * g.B + b) * g.wx + (ip - sx)) * g.wy + (jp - sy)];
      }
    }
    x[e] = acc;
  }
}

static int slice_geom(int64_t B, int64_t H, int64_t W, int64_t n, int64_t margin, SliceGeom* g) {
  if (B < 1 || H < 1 || W < 1 || n < 1 || margin < 0)
    return fail(kInvalidArgument, "invalid slice_patches geometry");
  const int64_t dx = (H - 2 * margin) / n, dy = (W - 2 * margin) / n;
  if (dx < 1 || dy < 1) return fail(kInvalidArgument, "patches larger than the grid");
  g->B = (int)B;
  g->H = (int)H;
  g->W = (int)W;
  g->n = (int)n;
  g->dx = (int)dx;
  g->dy = (int)dy;
  g->hx = (int)(dx / 2);
  g->hy = (int)(dy / 2);
  g->wx = (int)(2 * dx + 2 * margin);
  g->wy = (int)(2 * dy + 2 * margin);
  g->limx = g->H + 2 * g->hx - g->wx;
  g->limy = g->W + 2 * g->hy - g->wy;
  if (g->limx < 0 || g->limy < 0)
    return fail(kInvalidArgument, "slice window does not fit the padded grid");
  return kOk;
}

static int grid_for(int64_t n) {
  int64_t b = (n + PNT - 1) / PNT;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 16384));
}

int slice_patches(int dtype, const void* x, int64_t B, int64_t H, int64_t W, int64_t n,
                  int64_t margin, void* out, hipStream_t s) {
  SliceGeom g;
  if (int st = slice_geom(B, H, W, n, margin, &g)) return st;
  const int64_t total = (int64_t)n * n * B * g.wx * g.wy;
  if (dtype == 0)
    k_slice<float><<<grid_for(total), PNT, 0, s>>>(g, (const float*)x, (float*)out, total);
  else
    k_slice<double><<<grid_for(total), PNT, 0, s>>>(g, (const double*)x, (double*)out, total);
  return check_hip(hipGetLastError(), "slice_patches");
}

int slice_patches_adjoint(int dtype, const void* patches, int64_t B, int64_t H, int64_t W,
                          int64_t n, int64_t margin, void* x, hipStream_t s) {
  SliceGeom g;
  if (int st = slice_geom(B, H, W, n, margin, &g)) return st;
  const int64_t total = B * H * W;
  if (dtype == 0)
    k_slice_adj<float><<<grid_for(total), PNT, 0, s>>>(g, (const float*)patches, (float*)x,
                                                       total);
  else
    k_slice_adj<double><<<grid_for(total), PNT, 0, s>>>(g, (const double*)patches, (double*)x,
                                                        total);
  return check_hip(hipGetLastError(), "slice_patches_adjoint");
}

// ------------------------------------------------------------ convolve
// Real arrays over two axes of lengths H x W: c = x_kb + i y_b is
// transformed with one complex 2-D FFT, the spectra of x and y are separated
// with Z(-k), multiplied, and transformed back.  Batch b uses kernel kb = b /
// (nby / nbx).  Each axis runs at FFT length LH (LW): its own length when it
// has a compiled plan, else the smallest compiled length >= 2 n - 1 with the
// arrays zero-padded.  The size-L circular convolution of the padded arrays
// is then their linear convolution g (support 2 n - 1), and the size-n
// circular convolution the reference computes is the fold g[u] + g[u + n]:
// the row pass folds the columns after its inverse row FFT, the inverse
// column pass folds the rows (the fold commutes with the transform along the
// other axis).
struct ConvGeom {
  int H, W;      // array lengths of the two axes
  int LH, LW;    // their FFT lengths
  int nbx, nby;
  double scale;  // dvol / (LH LW)
};

template <typename T, int N>
struct ConvTile {
  static constexpr int V = Tile<T, N>::V < N ? Tile<T, N>::V : N;
};

// Column FFTs of length N = LH over the W columns, V per work-group.  PASS 0:
// forward FFT of x + i y (rows >= H zero) into the work planes (nby, LH, W);
// PASS 1: inverse FFT of the work planes, fold rows r + H onto r, real part
// times scale into out.  grid (ceil(W / V), nby).
template <typename T, int N, int PASS>
__global__ __launch_bounds__(PNT) void k_conv_cols(ConvGeom g, const T* __restrict__ x,
                                                   const T* __restrict__ y,
                                                   const cplx<T>* __restrict__ twg,
                                                   cplx<T>* __restrict__ Tw, T* __restrict__ out) {
  constexpr int V = ConvTile<T, N>::V;
  constexpr int VS = odd_stride(N);
  __shared__ cplx<T> buf[V * VS];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int kb = b / (g.nby / g.nbx);
  const int c0 = blockIdx.x * V;
  const int64_t plane = (int64_t)g.H * g.W;
  const int64_t wplane = (int64_t)N * g.W;
  load_twiddles(tw, twg, N, tid, PNT);
  for (int e = tid; e < V * N; e += PNT) {
    const int r = e / V, v = e - (e / V) * V;
    const int c = c0 + v;
    cplx<T> val = {T(0), T(0)};
    if (c < g.W) {
      const int64_t o = (int64_t)r * g.W + c;
      if (PASS == 1)
        val = Tw[b * wplane + o];
      else if (r < g.H)
        val = cplx<T>{x[kb * plane + o], y[b * plane + o]};
    }
    buf[v * VS + r] = val;
  }
  __syncthreads();
  fft_lds<T, N, V, VS, PNT, PASS == 1, kFewVectorPlan<N, V, PNT>>(buf, tw, tid);
  for (int e = tid; e < V * N; e += PNT) {
    const int r = e / V, v = e - (e / V) * V;
    const int c = c0 + v;
    if (c >= g.W) continue;
    const int64_t o = (int64_t)r * g.W + c;
    if (PASS == 0) {
      Tw[b * wplane + o] = buf[v * VS + r];
    } else if (r < g.H) {
      T s = buf[v * VS + r].re;
      if (N != g.H && r + g.H < N) s += buf[v * VS + r + g.H].re;
      out[b * plane + o] = T(g.scale) * s;
    }
  }
}

// Row pair (r, -r) of length N = LW: forward row FFTs (columns >= W zero),
// spectrum product X Y, inverse row FFTs, columns c + W folded onto c, the W
// folded columns written back.  grid (LH / 2 + 1, nby).
template <typename T, int N>
__global__ __launch_bounds__(PNT) void k_conv_rows(ConvGeom g, const cplx<T>* __restrict__ twg,
                                                   cplx<T>* __restrict__ Tw) {
  constexpr int VS = odd_stride(N);
  __shared__ cplx<T> buf[2 * VS];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x, r1 = (g.LH - r0) % g.LH;
  cplx<T>* base = Tw + (int64_t)blockIdx.y * g.LH * g.W;
  load_twiddles(tw, twg, N, tid, PNT);
  for (int e = tid; e < 2 * N; e += PNT) {
    const int v = e / N, k = e - (e / N) * N;
    buf[v * VS + k] = k < g.W ? base[(int64_t)(v ? r1 : r0) * g.W + k] : cplx<T>{T(0), T(0)};
  }
  __syncthreads();
  fft_lds<T, N, 2, VS, PNT, false, kFewVectorPlan<N, 2, PNT>>(buf, tw, tid);
  constexpr int EPT = (2 * N + PNT - 1) / PNT;
  cplx<T> prod[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = tid + u * PNT;
    if (e >= 2 * N) break;
    const int v = e / N, k = e - (e / N) * N;
    const cplx<T> z = buf[v * VS + k];
    const cplx<T> zm = buf[(1 - v) * VS + (k == 0 ? 0 : N - k)];
    const cplx<T> X = {T(0.5) * (z.re + zm.re), T(0.5) * (z.im - zm.im)};
    const cplx<T> Y = {T(0.5) * (z.im + zm.im), T(0.5) * (zm.re - z.re)};
    prod[u] = cmul(X, Y);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = tid + u * PNT;
    if (e >= 2 * N) break;
    buf[(e / N) * VS + e - (e / N) * N] = prod[u];
  }
  __syncthreads();
  fft_lds<T, N, 2, VS, PNT, true, kFewVectorPlan<N, 2, PNT>>(buf, tw, tid);
  const int nrows = r0 == r1 ? 1 : 2;
  for (int e = tid; e < nrows * g.W; e += PNT) {
    const int v = e / g.W, k = e - v * g.W;
    cplx<T> z = buf[v * VS + k];
    if (N != g.W && k + g.W < N) z = cadd(z, buf[v * VS + k + g.W]);
    base[(int64_t)(v ? r1 : r0) * g.W + k] = z;
  }
}

#define JUBIK_CONV_LENGTHS(X) \
  X(4) X(8) X(16) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(2048)

int convolve_fft_length(int64_t n) {
  if (n < 1) return 0;
  int best = 0;
#define JUBIK_CASE(l) \
  if (n == l) return l; \
  if (best == 0 && l >= 2 * n - 1) best = l;
  JUBIK_CONV_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
  return best;
}

template <typename T>
static void conv_cols(const ConvGeom& g, int pass, const void* x, const void* y, const void* tw,
                      void* work, void* out, hipStream_t s) {
  using C = cplx<T>;
  switch (g.LH) {
#define JUBIK_CASE(l)                                                                        \
  case l: {                                                                                  \
    constexpr int V = ConvTile<T, l>::V;                                                     \
    const dim3 gc((g.W + V - 1) / V, g.nby);                                                 \
    if (pass == 0)                                                                           \
      k_conv_cols<T, l, 0><<<gc, PNT, 0, s>>>(g, (const T*)x, (const T*)y, (const C*)tw,     \
                                               (C*)work, nullptr);                           \
    else                                                                                     \
      k_conv_cols<T, l, 1><<<gc, PNT, 0, s>>>(g, nullptr, nullptr, (const C*)tw, (C*)work,   \
                                               (T*)out);                                     \
    break;                                                                                   \
  }
    JUBIK_CONV_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    default:
      break;
  }
}

template <typename T>
static void conv_rows(const ConvGeom& g, const void* tw, void* work, hipStream_t s) {
  using C = cplx<T>;
  const dim3 gr(g.LH / 2 + 1, g.nby);
  switch (g.LW) {
#define JUBIK_CASE(l)                                                                   \
  case l:                                                                               \
    k_conv_rows<T, l><<<gr, PNT, 0, s>>>(g, (const C*)tw, (C*)work);                    \
    break;
    JUBIK_CONV_LENGTHS(JUBIK_CASE)
#undef JUBIK_CASE
    default:
      break;
  }
}

int convolve2d(int dtype, const void* x, const void* y, int64_t nbx, int64_t nby, int64_t H,
               int64_t W, double dvol, const void* twh, const void* tww, void* work, void* out,
               hipStream_t s) {
  if (nbx < 1 || nby < 1 || nby % nbx) return fail(kInvalidArgument, "batch sizes do not broadcast");
  const int LH = convolve_fft_length(H), LW = convolve_fft_length(W);
  if (!LH || !LW)
    return fail(kUnsupported, "convolve axis of length " + std::to_string(!LH ? H : W) +
                                  " exceeds 2048");
  ConvGeom g{(int)H, (int)W, LH, LW, (int)nbx, (int)nby, dvol / ((double)LH * (double)LW)};
  if (dtype == 0) {
    conv_cols<float>(g, 0, x, y, twh, work, out, s);
    conv_rows<float>(g, tww, work, s);
    conv_cols<float>(g, 1, x, y, twh, work, out, s);
  } else {
    conv_cols<double>(g, 0, x, y, twh, work, out, s);
    conv_rows<double>(g, tww, work, s);
    conv_cols<double>(g, 1, x, y, twh, work, out, s);
  }
  return check_hip(hipGetLastError(), "convolve");
}

}  // namespace jubik
