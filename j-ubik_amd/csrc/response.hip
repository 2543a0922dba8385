// Exposure, read-out and Poisson likelihood kernels
// (jubik/response.py:12-112; the NLL formula of nifty.re's Poissonian as
// documented at demos/x-ray-imaging.py:312-313).  Reductions run in FP64 in
// a fixed order: per-block partial sums, then one block sums the partials.
#include <mutex>
#include <string>
#include <type_traits>

#include "response.hpp"

namespace jubik {

constexpr int RNT = 256;
constexpr int RBLOCKS = 1024;  // grid of the reducing kernels (partials per call)

template <typename T>
__global__ __launch_bounds__(RNT) void k_exposure(const T* __restrict__ expo, int64_t n_e,
                                                  const T* __restrict__ x, int64_t n_x,
                                                  T* __restrict__ out, int64_t n_out) {
  for (int64_t i = blockIdx.x * (int64_t)RNT + threadIdx.x; i < n_out;
       i += (int64_t)gridDim.x * RNT)
    out[i] = expo[n_e == n_out ? i : i % n_e] * x[n_x == n_out ? i : i % n_x];
}

template <typename T>
__global__ __launch_bounds__(RNT) void k_gather(const T* __restrict__ x,
                                                const int64_t* __restrict__ idx,
                                                T* __restrict__ out, int64_t nnz) {
  for (int64_t k = blockIdx.x * (int64_t)RNT + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * RNT)
    out[k] = x[idx[k]];
}

template <typename T>
__global__ __launch_bounds__(RNT) void k_fill(T* __restrict__ out, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)RNT + threadIdx.x; k < n; k += (int64_t)gridDim.x * RNT)
    out[k] = T(0);
}

template <typename T>
__global__ __launch_bounds__(RNT) void k_scatter(const T* __restrict__ v,
                                                 const int64_t* __restrict__ idx,
                                                 T* __restrict__ out, int64_t nnz) {
  for (int64_t k = blockIdx.x * (int64_t)RNT + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * RNT)
    out[idx[k]] = v[k];
}

// partial[b] = sum over this block's elements of (lam - d log lam);
// grad = 1 - d / lam.
template <typename T>
__global__ __launch_bounds__(RNT) void k_poisson(const T* __restrict__ lam,
                                                 const T* __restrict__ d, int64_t n,
                                                 T* __restrict__ grad,
                                                 double* __restrict__ partial) {
  __shared__ double red[RNT / 64];
  double acc = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)RNT + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * RNT) {
    const T l = lam[k], dk = d[k];
    acc += (double)l - (double)dk * log((double)l);
    if (grad) grad[k] = T(1) - dk / l;
  }
  acc = block_sum<RNT>(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Metric of the Poisson likelihood in data space, diag(1/lam) (SQRT = 0),
// and its left square root diag(lam^-1/2) (SQRT = 1), applied to t.
template <typename T, bool SQRT>
__global__ __launch_bounds__(RNT) void k_poisson_metric(const T* __restrict__ lam,
                                                        const T* __restrict__ t, int64_t n,
                                                        T* __restrict__ out) {
  for (int64_t k = blockIdx.x * (int64_t)RNT + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * RNT)
    out[k] = SQRT ? t[k] / sqrt(lam[k]) : t[k] / lam[k];
}

// out[0] = sum of partial[0 .. n) in a fixed order.
__global__ __launch_bounds__(RNT) void k_reduce(const double* __restrict__ partial, int64_t n,
                                                double* __restrict__ out) {
  __shared__ double red[RNT / 64];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += RNT) acc += partial[k];
  acc = block_sum<RNT>(acc, red);
  if (threadIdx.x == 0) out[0] = acc;
}

// Cotangent of the metric's left square root J^T diag(lam^-1/2) at the base
// point: cot[s, i] = expo[i] * noise[s, i] / sqrt(lam[i]) where counts[i] >= 0,
// else 0; i runs over the n = J H W pixels of one sample, s over S samples.
template <typename T>
__global__ __launch_bounds__(RNT) void k_sqrt_metric_cot(const T* __restrict__ expo,
                                                         const int* __restrict__ counts,
                                                         const T* __restrict__ lam,
                                                         const T* __restrict__ noise,
                                                         T* __restrict__ cot, int64_t n,
                                                         int64_t total) {
  for (int64_t k = blockIdx.x * (int64_t)RNT + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * RNT) {
    const int64_t i = k % n;
    cot[k] = counts[i] >= 0 ? expo[i] * noise[k] / sqrt(lam[i]) : T(0);
  }
}

// ---------------------------------------------------------------- Poisson
// Counter-based uniforms: the SplitMix64 output function (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
// applied to seed + counter * golden gamma; 53-bit doubles in (0, 1).
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// One Poisson(lam) draw from the uniforms of stream ``base``.  lam < 12:
// inversion by sequential search of the CDF.  lam >= 12: the transformed
// rejection method with squeeze (PTRS) of W. Hormann, "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12 (1993) 39-45.  Both loops are bounded; the
// caps are never reached in practice (probability below 1e-40).
__device__ __forceinline__ int poisson_draw(double lam, uint64_t seed, uint64_t base) {
  if (!(lam > 0.0)) return 0;
  if (lam < 12.0) {
    const double u = uniform01(seed, base);
    double p = exp(-lam), F = p;
    int k = 0;
    while (u > F && k < 200) {
      ++k;
      p *= lam / k;
      F += p;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int it = 0; it < 128; ++it) {
    const double U = uniform01(seed, base + 2 * it) - 0.5;
    const double V = uniform01(seed, base + 2 * it + 1);
    const double us = 0.5 - fabs(U);
    const double kf = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (int)kf;
    if (kf < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + kf * loglam - lgamma(kf + 1.0))
      return (int)kf;
  }
  return (int)floor(lam + 0.5);
}

// counts[i] ~ Poisson(lam[i]); element i uses the uniform stream starting at
// counter 256 i, so a draw depends only on (seed, i, lam[i]).
template <typename T>
__global__ __launch_bounds__(RNT) void k_poisson_draw(const T* __restrict__ lam, int64_t n,
                                                      uint64_t seed, int* __restrict__ counts) {
  for (int64_t i = blockIdx.x * (int64_t)RNT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RNT)
    counts[i] = poisson_draw((double)lam[i], seed, (uint64_t)i * 256u);
}

static int grid_for(int64_t n) {
  int64_t b = (n + RNT - 1) / RNT;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

template <typename T>
static int exposure_t(const void* expo, int64_t n_e, const void* x, int64_t n_x, void* out,
                      int64_t n_out, hipStream_t s) {
  k_exposure<T><<<grid_for(n_out), RNT, 0, s>>>((const T*)expo, n_e, (const T*)x, n_x, (T*)out,
                                                n_out);
  return check_hip(hipGetLastError(), "exposure");
}

int exposure_apply(int dtype, const void* expo, int64_t n_e, const void* x, int64_t n_x,
                   void* out, int64_t n_out, hipStream_t s) {
  if (n_out <= 0) return kOk;
  if (dtype == 0) return exposure_t<float>(expo, n_e, x, n_x, out, n_out, s);
  return exposure_t<double>(expo, n_e, x, n_x, out, n_out, s);
}

int readout_gather(int dtype, const void* x, const int64_t* idx, int64_t nnz, void* out,
                   hipStream_t s) {
  if (nnz <= 0) return kOk;
  if (dtype == 0)
    k_gather<float><<<grid_for(nnz), RNT, 0, s>>>((const float*)x, idx, (float*)out, nnz);
  else
    k_gather<double><<<grid_for(nnz), RNT, 0, s>>>((const double*)x, idx, (double*)out, nnz);
  return check_hip(hipGetLastError(), "readout_gather");
}

int readout_scatter(int dtype, const void* v, const int64_t* idx, int64_t nnz, int64_t n,
                    void* out, hipStream_t s) {
  if (n <= 0) return kOk;
  if (dtype == 0) {
    k_fill<float><<<grid_for(n), RNT, 0, s>>>((float*)out, n);
    if (nnz > 0)
      k_scatter<float><<<grid_for(nnz), RNT, 0, s>>>((const float*)v, idx, (float*)out, nnz);
  } else {
    k_fill<double><<<grid_for(n), RNT, 0, s>>>((double*)out, n);
    if (nnz > 0)
      k_scatter<double><<<grid_for(nnz), RNT, 0, s>>>((const double*)v, idx, (double*)out, nnz);
  }
  return check_hip(hipGetLastError(), "readout_scatter");
}

int sqrt_metric_cot(int dtype, const void* expo, const int* counts, const void* lam,
                    const void* noise, void* cot, int64_t n, int S, hipStream_t s) {
  const int64_t total = n * S;
  if (total <= 0) return kOk;
  if (dtype == 0)
    k_sqrt_metric_cot<float><<<grid_for(total), RNT, 0, s>>>(
        (const float*)expo, counts, (const float*)lam, (const float*)noise, (float*)cot, n, total);
  else
    k_sqrt_metric_cot<double><<<grid_for(total), RNT, 0, s>>>(
        (const double*)expo, counts, (const double*)lam, (const double*)noise, (double*)cot, n,
        total);
  return check_hip(hipGetLastError(), "sqrt_metric_cot");
}

// geoVI: per read-out pixel geo = sqrt(lam_x lam_p) (the metric scaling of
// the cross operator J(p)^T D_p D_x J(x)) and eta = 2 (sqrt lam_x - sqrt
// lam_p) (the change of the Poisson transformation 2 sqrt lam); masked
// pixels get geo = 1, eta = 0.
template <typename T>
__global__ __launch_bounds__(RNT) void k_geo_prep(const int* __restrict__ counts,
                                                  const T* __restrict__ lam_p,
                                                  const T* __restrict__ lam_x,
                                                  T* __restrict__ geo, T* __restrict__ eta,
                                                  int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)RNT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RNT) {
    const bool ok = counts[i] >= 0;
    const T lp = lam_p[i], lx = lam_x[i];
    geo[i] = ok ? sqrt(lx * lp) : T(1);
    eta[i] = ok ? T(2) * (sqrt(lx) - sqrt(lp)) : T(0);
  }
}

int geo_prep(int dtype, const int* counts, const void* lam_p, const void* lam_x, void* geo,
             void* eta, int64_t n, hipStream_t s) {
  if (n <= 0) return kOk;
  if (dtype == 0)
    k_geo_prep<float><<<grid_for(n), RNT, 0, s>>>(counts, (const float*)lam_p,
                                                  (const float*)lam_x, (float*)geo, (float*)eta,
                                                  n);
  else
    k_geo_prep<double><<<grid_for(n), RNT, 0, s>>>(counts, (const double*)lam_p,
                                                   (const double*)lam_x, (double*)geo,
                                                   (double*)eta, n);
  return check_hip(hipGetLastError(), "geo_prep");
}

int poisson_sample(int dtype, const void* lam, int64_t n, uint64_t seed, int* counts,
                   hipStream_t s) {
  if (n <= 0) return kOk;
  if (dtype == 0)
    k_poisson_draw<float><<<grid_for(n), RNT, 0, s>>>((const float*)lam, n, seed, counts);
  else
    k_poisson_draw<double><<<grid_for(n), RNT, 0, s>>>((const double*)lam, n, seed, counts);
  return check_hip(hipGetLastError(), "poisson_sample");
}

int poisson_metric(int dtype, const void* lam, const void* t, int64_t n, int sqrt_form,
                   void* out, hipStream_t s) {
  if (n <= 0) return kOk;
  auto go = [&](auto tag, auto sq) {
    using T = decltype(tag);
    k_poisson_metric<T, decltype(sq)::value><<<grid_for(n), RNT, 0, s>>>(
        (const T*)lam, (const T*)t, n, (T*)out);
  };
  if (dtype == 0) {
    if (sqrt_form) go(float{}, std::true_type{});
    else go(float{}, std::false_type{});
  } else {
    if (sqrt_form) go(double{}, std::true_type{});
    else go(double{}, std::false_type{});
  }
  return check_hip(hipGetLastError(), "poisson_metric");
}

// Blocks of the Poisson reduction for n elements: one partial per block.
static int poisson_blocks(int64_t n) {
  return (int)std::min<int64_t>(std::max<int64_t>((n + RNT - 1) / RNT, 1), RBLOCKS);
}

int poisson_nll_grad(int dtype, const void* lam, const void* d, int64_t n, double* nll,
                     void* grad, double* work, hipStream_t s) {
  const int blocks = poisson_blocks(n);
  if (dtype == 0)
    k_poisson<float><<<blocks, RNT, 0, s>>>((const float*)lam, (const float*)d, n, (float*)grad,
                                            work);
  else
    k_poisson<double><<<blocks, RNT, 0, s>>>((const double*)lam, (const double*)d, n,
                                             (double*)grad, work);
  k_reduce<<<1, RNT, 0, s>>>(work, blocks, nll);
  return check_hip(hipGetLastError(), "poisson_nll_grad");
}

// ---------------------------------------------------------------- no-allocation variant
// The partials of the Poisson reduction in a file-scope device array, for
// callers that cannot allocate (graph capture, serving loops): one copy per
// GPU, held until the process ends.  Host code never passes ``scratch``
// itself to a kernel: its device address comes from hipGetSymbolAddress,
// once per device, with that device current.
static_assert(RBLOCKS <= kPoissonWorkDoubles, "the reduction's partials fit the scratch");
static __device__ __align__(256) double scratch[kPoissonWorkDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of this device's scratch array (looked up on first use).
static int scratch_address(double** out) {
  *out = nullptr;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return check_hip(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(kUnsupported, "device index " + std::to_string(dev) + " beyond " +
                                  std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

int poisson_nll_grad_noalloc(int dtype, const void* lam, const void* d, int64_t n, double* nll,
                             void* grad, hipStream_t s) {
  const int blocks = poisson_blocks(n);
  if (blocks > kPoissonWorkDoubles)
    return fail(kInvalidArgument, "the reduction needs " + std::to_string(blocks) +
                                      " partials; the scratch holds " +
                                      std::to_string(kPoissonWorkDoubles));
  double* work = nullptr;
  const int st = scratch_address(&work);
  if (st) return st;
  // Every partial the reduction reads is written by this call's first
  // kernel, so the scratch needs no clearing.
  return poisson_nll_grad(dtype, lam, d, n, nll, grad, work, s);
}

}  // namespace jubik
