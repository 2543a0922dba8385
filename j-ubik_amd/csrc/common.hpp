// Shared definitions of the jubik HIP library: status codes, error text,
// launch checking and the geometry of the linear-patching PSF operator.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "fft.hpp"

namespace jubik {

enum Status : int {
  kOk = 0,
  kInvalidArgument = -1,
  kUnsupported = -2,
  kRuntimeError = -3,
  kOutOfMemory = -4,
  kNotReady = -5,
  kDomain = -6,
};

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// Optional per-launch timing (jubik_timing_enable): HIP events recorded on
// the launch stream around each kernel (or only around the kernels of one
// name, jubik_timing_select), aggregated by name on report.
bool timing_enabled();
bool timing_wanted(const char* name);
void timing_record(const char* name, hipEvent_t start, hipEvent_t stop);
struct TimedLaunch {
  const char* name;
  hipStream_t stream;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(const char* n, hipStream_t s) : name(n), stream(s) {
    if (timing_wanted(n) && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, s);
    else
      a = b = nullptr;
  }
  ~TimedLaunch() {
    if (a && b) {
      (void)hipEventRecord(b, stream);
      timing_record(name, a, b);
    }
  }
};

#define JUBIK_HIP(call)                                  \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return check_hip(_e, #call);   \
  } while (0)

// Geometry of linpatch_convolve on one (H, W) grid with n patches per axis
// (jubik/convolve.py:66-145).  Windows are w = 2d wide with stride d; the
// slicing pads the grid by h = d / 2 zeros; every window is padded by the
// margin m to the FFT length L = w + 2m.  Window starts are
// min(i * d, lim), the clamping of jax.lax.dynamic_slice; the same starts
// serve the slicing of the image and the overlap-add canvas.
// Row stride of the half-spectrum arrays (R1, R2, kernel spectra): the L/2
// stored complex values (DC and Nyquist packed in slot 0) rounded up to a
// multiple of 16 (128 bytes in FP32), so the 16-column tiles of the column
// pass are single aligned cache lines, and to an odd number of such blocks:
// a column tile walking down the rows then spreads over the HBM channels
// (an even stride, e.g. 12 lines at L = 384, concentrates it on few of them
// and measured 10 % slower).  Padding blocks are never read or written.
__host__ __device__ constexpr int half_stride(int L) {
  const int q = (L / 2 + 15) / 16;
  return 16 * (q % 2 ? q : q + 1);
}

struct LinGeom {
  int H, W;       // sky grid
  int n;          // patches per axis
  int d;          // patch stride (equal on both axes)
  int h;          // zero padding of the slicing, d / 2
  int w;          // window 2d
  int m;          // margin
  int L;          // FFT length: w + 2m, or a longer compiled length (below)
  int L2;         // L / 2 stored half-spectrum columns (DC and Nyquist packed)
  // Window length Lw = w + 2m of the reference's circular convolution and
  // the margin mf = (L - w) / 2 of the FFT frame.  Normally L = Lw and
  // mf = m.  When w + 2m has no compiled plan the transforms run at a
  // compiled L >= max(Lw, 2w): with mf >= w / 2 a window's size-L circular
  // convolution is its linear convolution g (support [-w/2, 3w/2 - 1) in
  // window coordinates), and the size-Lw result is the fold
  // sum_j g[u + j Lw] for u in [-m, w + m), j in {-1, 0, 1}; the adjoint
  // reads the periodic extension of the Lw window ("fold mode").
  int Lw;
  int mf;
  int limx, limy; // clamp of the window starts: H + 2h - w, W + 2h - w
  int P;          // n * n
  int E;          // energies per module
  int M;          // modules
  int J;          // jobs M * E
  int x_shared;   // 1: the input image of job (mod, e) is x[e]
  int Ek;         // kernel energies: E, or 1 when one kernel serves all
  int S;          // samples per launch: jobs (s, mod, e) = s * J + mod * E + e
  int sky_shared; // adjoint into the GP: 1 = one sky (E, H, W) for every
                  // sample (metric at a base point), 0 = one sky per sample
};

__device__ __forceinline__ int start_x(const LinGeom& g, int i) { return min(i * g.d, g.limx); }
__device__ __forceinline__ int start_y(const LinGeom& g, int i) { return min(i * g.d, g.limy); }
// Spectrum index of (job, patch): kernels are laid out (M, P, Ek); every
// sample of a batch shares them.
__device__ __forceinline__ int64_t spec_index(const LinGeom& g, int job, int p) {
  const int jm = job % g.J;
  const int mod = jm / g.E, e = jm - mod * g.E;
  return ((int64_t)mod * g.P + p) * g.Ek + (g.Ek == 1 ? 0 : e);
}

// Index of the shared row spectra of job z of sample smp: the job, or for a
// forward signal shared by the modules the image (smp, energy).
__device__ __forceinline__ int r1_image(const LinGeom& g, int smp, int z, bool fwd) {
  return (fwd && g.x_shared) ? smp * g.E + z % g.E : smp * g.J + z;
}

// Tent weight t of a 2d window (jubik/convolve.py:18-25): numpy
// linspace(0, 1, d) then its mirror image; entry i of linspace is i * step
// with step = 1 / (d - 1) in double and the last entry exactly 1.  IEEE
// double division and multiply give the same values as the host table.
template <typename T>
__device__ __forceinline__ T tent_weight(int d, double step, int t) {
  const int i = t < d ? t : 2 * d - 1 - t;
  return (T)((d > 1 && i == d - 1) ? 1.0 : i * step);
}
__device__ __forceinline__ double tent_step(int d) { return d > 1 ? 1.0 / (d - 1) : 0.0; }

// Streaming access to workspace arrays that are written once and read once
// by the next pass: non-temporal loads and stores.
template <typename T>
__device__ __forceinline__ cplx<T> ld_stream(const cplx<T>* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return cplx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void st_stream(cplx<T>* p, cplx<T> c) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 v;
  v.x = c.re;
  v.y = c.im;
  __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
}

// Bounds-checked gathers through a raw buffer resource (stride 0, range
// checked on the byte offset): a load at an offset at or beyond ``bytes``
// returns zero in hardware.  Callers pass that offset for elements outside
// the array, so a gather needs neither a branch nor a select and all its
// loads are in flight together.  Word 3 is the gfx9-family default format.
// Offsets are 32-bit: gathered arrays stay below 2 GiB (linpatch images are
// at most 4096 pixels wide; the GP adjoint checks its planes on the host).
struct Gather {
  __amdgpu_buffer_rsrc_t r;
  int oob;  // an offset that reads zero
};
template <typename T>
__device__ __forceinline__ Gather make_gather(const T* p, int64_t count) {
  const int bytes = (int)(count * (int64_t)sizeof(T));
  return Gather{__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), 0, bytes, 0x00020000),
                bytes};
}
// Byte offset of element i, or the zero-reading offset when !ok.
template <typename T>
__device__ __forceinline__ int gather_off(const Gather& g, bool ok, int i) {
  return ok ? i * (int)sizeof(T) : g.oob;
}
template <typename T>
__device__ __forceinline__ T gather_load(const Gather& g, int off) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "gather element size");
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(g.r, off, 0, 0));
  else if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(g.r, off, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(g.r, off, 0, 0));
}

// Tile sizes: vectors per work-group so a tile of V vectors of length L at
// the odd LDS stride occupies at most 52 KiB (16 float vectors at L = 384).
#ifndef JUBIK_COL_TILE_BYTES
#define JUBIK_COL_TILE_BYTES 53248
#endif
#define JUBIK_ROW_TILE_BYTES 13312
__host__ __device__ constexpr int tile_vectors(int bytes, int L, int elem) {
  const int raw = bytes / ((L | 1) * elem);
  return raw >= 32 ? 32 : raw >= 16 ? 16 : raw >= 8 ? 8 : raw >= 4 ? 4 : raw >= 2 ? 2 : 1;
}
template <typename T, int L>
struct Tile {
  static constexpr int V = tile_vectors(JUBIK_COL_TILE_BYTES, L, (int)sizeof(cplx<T>));
};

// Row passes use a quarter of that LDS budget: their global traffic is
// contiguous rows whatever V is, and the smaller footprint raises residency
// (at L = 384, 4 vectors per tile measured faster than 8 or 16).
template <typename T, int L>
struct RowTile {
  static constexpr int V = tile_vectors(JUBIK_ROW_TILE_BYTES, L, (int)sizeof(cplx<T>));
};

// Tile of work-group b out of nb: the dispatcher places work-group b on XCD
// b mod 8, so consecutive tiles are given to the work-groups of one XCD and
// neighbouring tiles that share cache lines meet in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int kXcds = 8;
  if (nb % kXcds) return b;
  return (b % kXcds) * (nb / kXcds) + b / kXcds;
}

// Block-wide sum of one double per thread (NT a multiple of 64).
template <int NT>
__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  return s;  // valid in thread 0
}

}  // namespace jubik
