// Device-resident conjugate gradients (see cg.hpp).  Per-system state in
// device doubles: [0] rr = r.r, [1] bb = b.b, [2] pAp, [3] alpha, [4] beta,
// [5] active (1 / 0), [6] iterations, [7] energy 1/2 x.A.x - b.x, [8] rr of
// the update, [9] energy of the update.  Vector kernels run a (kCgBlocks,
// S) grid; each block reduces its strided share of a system into one FP64
// partial and k_cg_scalars sums the partials in a fixed order, so results
// are deterministic.
#include <vector>

#include "cg.hpp"

namespace jubik {

constexpr int CNT = 256;

int64_t cg_work_doubles(int S) { return (int64_t)S * kCgStateDoubles + 3LL * S * kCgBlocks; }

namespace {

template <typename T>
__device__ __forceinline__ T* field(const CgVec& v, int f, int s, int64_t n) {
  return static_cast<T*>(v.f[f]) + (int64_t)s * n;
}

// partial[k * S * B + s * B + b] for slot k
__device__ __forceinline__ double* partial_slot(double* work, int S, int k) {
  return work + (int64_t)S * kCgStateDoubles + (int64_t)k * S * kCgBlocks;
}

template <int NSLOT>
__device__ __forceinline__ void store_partials(double* work, int S, double (&acc)[NSLOT],
                                               double* red) {
#pragma unroll
  for (int k = 0; k < NSLOT; ++k) {
    const double v = block_sum<CNT>(acc[k], red);
    if (threadIdx.x == 0) partial_slot(work, S, k)[blockIdx.y * kCgBlocks + blockIdx.x] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(CNT) void k_cg_init(CgDims d, CgVec b, CgVec x, CgVec r, CgVec p,
                                                 double* work) {
  __shared__ double red[CNT / 64];
  const int s = blockIdx.y;
  double acc[1] = {0.0};
  for (int f = 0; f < d.nf; ++f) {
    const T* bf = field<T>(b, f, s, d.n);
    T* xf = field<T>(x, f, s, d.n);
    T* rf = field<T>(r, f, s, d.n);
    T* pf = field<T>(p, f, s, d.n);
    for (int64_t i = blockIdx.x * (int64_t)CNT + threadIdx.x; i < d.n;
         i += (int64_t)kCgBlocks * CNT) {
      const T v = bf[i];
      xf[i] = T(0);
      rf[i] = v;
      pf[i] = v;
      acc[0] += (double)v * (double)v;
    }
  }
  store_partials<1>(work, d.S, acc, red);
}

template <typename T>
__global__ __launch_bounds__(CNT) void k_cg_matvec(CgDims d, CgVec p, CgVec Mv, T c, CgVec Ap,
                                                   double* work) {
  __shared__ double red[CNT / 64];
  const int s = blockIdx.y;
  double acc[1] = {0.0};
  for (int f = 0; f < d.nf; ++f) {
    const T* pf = field<T>(p, f, s, d.n);
    const T* mf = field<T>(Mv, f, s, d.n);
    T* af = field<T>(Ap, f, s, d.n);
    for (int64_t i = blockIdx.x * (int64_t)CNT + threadIdx.x; i < d.n;
         i += (int64_t)kCgBlocks * CNT) {
      const T a = c * mf[i] + pf[i];
      af[i] = a;
      acc[0] += (double)pf[i] * (double)a;
    }
  }
  store_partials<1>(work, d.S, acc, red);
}

template <typename T>
__global__ __launch_bounds__(CNT) void k_cg_update(CgDims d, CgVec b, CgVec x, CgVec r, CgVec p,
                                                   CgVec Ap, double* work) {
  __shared__ double red[CNT / 64];
  const int s = blockIdx.y;
  const T alpha = T(work[s * kCgStateDoubles + 3]);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int f = 0; f < d.nf; ++f) {
    const T* bf = field<T>(b, f, s, d.n);
    T* xf = field<T>(x, f, s, d.n);
    T* rf = field<T>(r, f, s, d.n);
    const T* pf = field<T>(p, f, s, d.n);
    const T* af = field<T>(Ap, f, s, d.n);
    for (int64_t i = blockIdx.x * (int64_t)CNT + threadIdx.x; i < d.n;
         i += (int64_t)kCgBlocks * CNT) {
      const T xn = xf[i] + alpha * pf[i];
      const T rn = rf[i] - alpha * af[i];
      xf[i] = xn;
      rf[i] = rn;
      acc[0] += (double)rn * (double)rn;
      acc[1] += (double)xn * (double)bf[i];
      acc[2] += (double)xn * (double)rn;
    }
  }
  store_partials<3>(work, d.S, acc, red);
}

template <typename T>
__global__ __launch_bounds__(CNT) void k_cg_direction(CgDims d, CgVec r, CgVec p,
                                                      const double* work) {
  const int s = blockIdx.y;
  const T beta = T(work[s * kCgStateDoubles + 4]);
  for (int f = 0; f < d.nf; ++f) {
    const T* rf = field<T>(r, f, s, d.n);
    T* pf = field<T>(p, f, s, d.n);
    for (int64_t i = blockIdx.x * (int64_t)CNT + threadIdx.x; i < d.n;
         i += (int64_t)kCgBlocks * CNT)
      pf[i] = rf[i] + beta * pf[i];
  }
}

// One block per system: sums the partials of `nslot` slots in a fixed order
// and advances the scalar recurrences of `phase` (0 init, 1 after the
// matrix-vector product, 2 after the update).
__global__ __launch_bounds__(CNT) void k_cg_scalars(CgDims d, double* work, int phase) {
  __shared__ double red[CNT / 64];
  const int s = blockIdx.x;
  const int nslot = phase == 2 ? 3 : 1;
  double sum[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nslot; ++k) {
    const double* ps = partial_slot(work, d.S, k) + (int64_t)s * kCgBlocks;
    double a = 0.0;
    for (int i = threadIdx.x; i < kCgBlocks; i += CNT) a += ps[i];
    sum[k] = block_sum<CNT>(a, red);
  }
  if (threadIdx.x != 0) return;
  double* st = work + (int64_t)s * kCgStateDoubles;
  if (phase == 0) {
    st[0] = sum[0];
    st[1] = sum[0];
    st[2] = st[3] = st[4] = 0.0;
    st[5] = sum[0] > 0.0 ? 1.0 : 0.0;
    st[6] = 0.0;
    st[7] = 0.0;
  } else if (phase == 1) {
    const bool active = st[5] != 0.0;
    const double pAp = sum[0];
    st[2] = pAp;
    st[3] = (active && pAp > 0.0) ? st[0] / pAp : 0.0;
    if (active && !(pAp > 0.0)) st[5] = 0.0;  // breakdown: keep the last iterate
  } else {
    const bool active = st[5] != 0.0;
    const double rr_new = sum[0];
    const double e_new = -0.5 * (sum[1] + sum[2]);  // 1/2 x.A.x - b.x = -1/2 x.(b + r)
    st[8] = rr_new;
    st[9] = e_new;
    if (active) {
      st[6] += 1.0;
      bool done = rr_new <= d.rtol * d.rtol * st[1];
      if (d.absdelta > 0.0 && fabs(st[7] - e_new) <= d.absdelta) done = true;
      st[4] = st[0] > 0.0 ? rr_new / st[0] : 0.0;
      st[0] = rr_new;
      st[7] = e_new;
      if (done) st[5] = 0.0;
    } else {
      st[4] = 0.0;
    }
  }
}

__global__ void k_cg_finish(CgDims d, const double* work, double* info) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.S) return;
  const double* st = work + (int64_t)s * kCgStateDoubles;
  info[2 * s] = st[6];
  info[2 * s + 1] = st[1] > 0.0 ? sqrt(st[0] / st[1]) : 0.0;
}

}  // namespace

static int cg_status(const char* what) { return check_hip(hipGetLastError(), what); }

int cg_init(int dtype, const CgDims& d, CgVec b, CgVec x, CgVec r, CgVec p, double* work,
            hipStream_t s) {
  const dim3 gr(kCgBlocks, d.S);
  if (dtype == 0) k_cg_init<float><<<gr, CNT, 0, s>>>(d, b, x, r, p, work);
  else k_cg_init<double><<<gr, CNT, 0, s>>>(d, b, x, r, p, work);
  k_cg_scalars<<<d.S, CNT, 0, s>>>(d, work, 0);
  return cg_status("cg_init");
}

int cg_matvec(int dtype, const CgDims& d, CgVec p, CgVec Mv, double c, CgVec Ap, double* work,
              hipStream_t s) {
  const dim3 gr(kCgBlocks, d.S);
  if (dtype == 0) k_cg_matvec<float><<<gr, CNT, 0, s>>>(d, p, Mv, (float)c, Ap, work);
  else k_cg_matvec<double><<<gr, CNT, 0, s>>>(d, p, Mv, c, Ap, work);
  k_cg_scalars<<<d.S, CNT, 0, s>>>(d, work, 1);
  return cg_status("cg_matvec");
}

int cg_update(int dtype, const CgDims& d, CgVec b, CgVec x, CgVec r, CgVec p, CgVec Ap,
              double* work, hipStream_t s) {
  const dim3 gr(kCgBlocks, d.S);
  if (dtype == 0) k_cg_update<float><<<gr, CNT, 0, s>>>(d, b, x, r, p, Ap, work);
  else k_cg_update<double><<<gr, CNT, 0, s>>>(d, b, x, r, p, Ap, work);
  k_cg_scalars<<<d.S, CNT, 0, s>>>(d, work, 2);
  if (dtype == 0) k_cg_direction<float><<<gr, CNT, 0, s>>>(d, r, p, work);
  else k_cg_direction<double><<<gr, CNT, 0, s>>>(d, r, p, work);
  return cg_status("cg_update");
}

int cg_finish(const CgDims& d, const double* work, double* info, hipStream_t s) {
  k_cg_finish<<<(d.S + 63) / 64, 64, 0, s>>>(d, work, info);
  return cg_status("cg_finish");
}

int cg_active_count(const CgDims& d, const double* work, hipStream_t s, int* count) {
  std::vector<double> st((size_t)d.S * kCgStateDoubles);
  JUBIK_HIP(hipMemcpyAsync(st.data(), work, st.size() * sizeof(double), hipMemcpyDeviceToHost,
                           s));
  JUBIK_HIP(hipStreamSynchronize(s));
  int c = 0;
  for (int k = 0; k < d.S; ++k) c += st[(size_t)k * kCgStateDoubles + 5] != 0.0;
  *count = c;
  return kOk;
}

template <typename T>
__global__ __launch_bounds__(CNT) void k_vec_sum3(int64_t n, const T* __restrict__ a,
                                                 const T* __restrict__ b,
                                                 const T* __restrict__ c, T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)CNT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * CNT)
    out[i] = a[i] + b[i] + c[i];
}

int vec_sum3(int dtype, int64_t n, const void* a, const void* b, const void* c, void* out,
             hipStream_t s) {
  if (n <= 0) return kOk;
  const unsigned grid = (unsigned)((n + CNT - 1) / CNT < 4096 ? (n + CNT - 1) / CNT : 4096);
  if (dtype == 0)
    k_vec_sum3<float><<<grid, CNT, 0, s>>>(n, (const float*)a, (const float*)b,
                                           (const float*)c, (float*)out);
  else
    k_vec_sum3<double><<<grid, CNT, 0, s>>>(n, (const double*)a, (const double*)b,
                                            (const double*)c, (double*)out);
  return check_hip(hipGetLastError(), "vec_sum3");
}

}  // namespace jubik
