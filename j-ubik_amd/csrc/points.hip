// Point-source sky component: per pixel
//   p_e = I(xi) * exp(alpha * dnu_e),  I = invgamma_prior(a, q)(xi),
//   alpha = alpha_mean + alpha_std * xi_alpha
// (jubik/sky_models.py:376-441, the SkyModel point sources added to the
// diffuse sky at 202; spectral_product_mf_point_sources.py:99-287,
// MultiFrequencyInvGammaSky, without spectral deviations).
//
// The inverse-gamma prior maps a standard normal xi to the inverse-gamma
// variable of shape a and scale q with the same quantile: I = q / y with
// Q(a, y) = Phi(xi), Q the regularised upper incomplete gamma function (if
// Y ~ Gamma(a, 1) then P(q / Y <= x) = Q(a, q / x)).  y is found by Newton
// steps in log y on log P(a, y) = log Phi(-xi) (small y) or log Q(a, y) =
// log Phi(xi) (large y), started from the Wilson-Hilferty approximation of
// the Gamma quantile (whose standard normal quantile is exactly -xi) or, for
// small y or a < 1, from the leading term y^a / Gamma(a + 1) of P.  P and Q
// come from their power series (y < a + 1) or the continued fraction of Q
// (modified Lentz), in double precision.  The derivative is closed-form:
//   d log I / d xi = phi(xi) Gamma(a) / (y^a e^-y).
#include "points.hpp"

namespace jubik {

namespace {

constexpr int PSNT = 256;

// log P(a, y) and log Q(a, y), the regularised lower and upper incomplete
// gamma functions, y > 0.
__host__ __device__ void log_pq(double a, double y, double& lp, double& lq) {
  const double lpre = -y + a * log(y) - lgamma(a);  // log(y^a e^-y / Gamma(a))
  if (y < a + 1.0) {
    // P = pre * sum_n y^n / (a (a + 1) ... (a + n))
    double term = 1.0 / a, sum = term;
    for (int n = 1; n < 1000; ++n) {
      term *= y / (a + n);
      sum += term;
      if (term < sum * 1e-17) break;
    }
    lp = lpre + log(sum);
    lq = log1p(-exp(lp));
  } else {
    // Q = pre / (y + 1 - a - 1 (1 - a) / (y + 3 - a - 2 (2 - a) / ...))
    constexpr double tiny = 1e-300;
    double b = y + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 1000; ++i) {
      const double an = -i * (i - a);
      b += 2.0;
      d = an * d + b;
      if (fabs(d) < tiny) d = tiny;
      c = b + an / c;
      if (fabs(c) < tiny) c = tiny;
      d = 1.0 / d;
      const double del = d * c;
      h *= del;
      if (fabs(del - 1.0) < 1e-16) break;
    }
    lq = lpre + log(h);
    lp = log1p(-exp(lq));
  }
}

// log Phi(x) for the standard normal distribution function.
__host__ __device__ double log_ndtr(double x) { return log(0.5 * erfc(-x * 0.70710678118654752440)); }

// y with Q(a, y) = Phi(xi); xi is clamped to [-37, 37] (Phi stays normal).
__host__ __device__ double invgamma_y(double a, double xi) {
  xi = fmin(fmax(xi, -37.0), 37.0);
  const double lp_t = log_ndtr(-xi);  // target log P
  const double lq_t = log_ndtr(xi);   // target log Q
  const bool use_p = xi > 0.0;         // the smaller of the two is the accurate one
  // starting point
  double t;  // log y
  const double z = -xi;                // standard normal quantile of P's target
  const double lsmall = (lp_t + lgamma(a + 1.0)) / a;  // P ~ y^a / Gamma(a + 1)
  if (a >= 1.0) {
    const double c = 1.0 - 1.0 / (9.0 * a) + z / (3.0 * sqrt(a));
    t = c > 0.05 ? log(a) + 3.0 * log(c) : lsmall;
  } else {
    // P(a, y) ~ tt (y)^a below y = 1 and ~ 1 - (1 - tt) e^(1 - y) above,
    // tt = 1 - a (0.253 + 0.12 a) (the upper branch from Q, in logs)
    const double tt = 1.0 - a * (0.253 + 0.12 * a);
    t = lp_t < log(tt) ? (lp_t - log(tt)) / a : log(1.0 - lq_t + log(1.0 - tt));
  }
  for (int it = 0; it < 100; ++it) {
    const double y = exp(t);
    double lp, lq;
    log_pq(a, y, lp, lq);
    const double lf = a * t - y - lgamma(a);  // log(y f(y)), f the Gamma density
    double F, dF;                             // F(t) and dF/dt
    if (use_p) {
      F = lp - lp_t;
      dF = exp(lf - lp);
    } else {
      F = lq - lq_t;
      dF = -exp(lf - lq);
    }
    double step = F / dF;
    step = fmin(fmax(step, -2.0), 2.0);
    t -= step;
    if (fabs(step) < 1e-14 * fmax(1.0, fabs(t))) break;
  }
  return exp(t);
}

// log I, d log I / d xi and alpha per pixel.
template <typename T>
__global__ __launch_bounds__(PSNT) void k_ps_prepare(int64_t n, const T* __restrict__ xi,
                                                    const T* __restrict__ xa, double a, double q,
                                                    double amean, double astd, T* __restrict__ w) {
  const int64_t i = blockIdx.x * (int64_t)PSNT + threadIdx.x;
  if (i >= n) return;
  const double x = (double)xi[i];
  const double y = invgamma_y(a, x);
  const double xc = fmin(fmax(x, -37.0), 37.0);
  // d log I / d xi = phi(xi) Gamma(a) / (y^a e^-y)
  const double dl = exp(-0.5 * xc * xc - 0.91893853320467274178 + lgamma(a) - a * log(y) + y);
  w[i] = (T)(log(q) - log(y));
  w[n + i] = (T)dl;
  w[2 * n + i] = xa ? (T)(amean + astd * (double)xa[i]) : T(0);
}

// skyt[s, e] = sky[s, e] + exp(log I[s] + alpha[s] * dnu[e]).
template <typename T>
__global__ __launch_bounds__(PSNT) void k_ps_add(int S, int E, int64_t n, const T* __restrict__ sky,
                                                const T* __restrict__ w,
                                                const T* __restrict__ dnu,
                                                T* __restrict__ skyt) {
  const int64_t i = blockIdx.x * (int64_t)PSNT + threadIdx.x;
  if (i >= n) return;
  const int s = blockIdx.y;
  const T* ws = w + (int64_t)s * 3 * n;
  const T li = ws[i], al = ws[2 * n + i];
  for (int e = 0; e < E; ++e) {
    const int64_t o = ((int64_t)s * E + e) * n + i;
    skyt[o] = sky[o] + exp(li + (dnu ? al * dnu[e] : T(0)));
  }
}

// g_xi = d log I / d xi * sum_e p_e ct_e, g_xa = astd * sum_e p_e ct_e dnu_e,
// ct the cotangent of the total sky.
template <typename T>
__global__ __launch_bounds__(PSNT) void k_ps_grad(int S, int E, int64_t n, const T* __restrict__ ct,
                                                 const T* __restrict__ w,
                                                 const T* __restrict__ dnu, double astd,
                                                 T* __restrict__ g_xi, T* __restrict__ g_xa) {
  const int64_t i = blockIdx.x * (int64_t)PSNT + threadIdx.x;
  if (i >= n) return;
  const int s = blockIdx.y;
  const T* ws = w + (int64_t)s * 3 * n;
  const T li = ws[i], dl = ws[n + i], al = ws[2 * n + i];
  T a0 = T(0), a1 = T(0);
  for (int e = 0; e < E; ++e) {
    const T de = dnu ? dnu[e] : T(0);
    const T pc = exp(li + al * de) * ct[((int64_t)s * E + e) * n + i];
    a0 += pc;
    a1 += pc * de;
  }
  g_xi[(int64_t)s * n + i] = dl * a0;
  if (g_xa) g_xa[(int64_t)s * n + i] = T(astd) * a1;
}

template <typename T>
__global__ __launch_bounds__(PSNT) void k_invgamma(int64_t n, const T* __restrict__ xi, double a,
                                                  double q, T* __restrict__ logi,
                                                  T* __restrict__ dlogi) {
  const int64_t i = blockIdx.x * (int64_t)PSNT + threadIdx.x;
  if (i >= n) return;
  const double x = (double)xi[i];
  const double y = invgamma_y(a, x);
  const double xc = fmin(fmax(x, -37.0), 37.0);
  logi[i] = (T)(log(q) - log(y));
  if (dlogi)
    dlogi[i] = (T)exp(-0.5 * xc * xc - 0.91893853320467274178 + lgamma(a) - a * log(y) + y);
}

unsigned blocks(int64_t n) { return (unsigned)((n + PSNT - 1) / PSNT); }

}  // namespace

int invgamma_prior(int dtype, const void* xi, int64_t n, double a, double q, void* logi,
                   void* dlogi, hipStream_t s) {
  if (n <= 0) return kOk;
  if (dtype == 0)
    k_invgamma<float><<<blocks(n), PSNT, 0, s>>>(n, (const float*)xi, a, q, (float*)logi,
                                                 (float*)dlogi);
  else
    k_invgamma<double><<<blocks(n), PSNT, 0, s>>>(n, (const double*)xi, a, q, (double*)logi,
                                                  (double*)dlogi);
  return check_hip(hipGetLastError(), "invgamma_prior");
}

int ps_prepare(int dtype, const PointSources& ps, int S, int64_t n, const void* xi,
               const void* xa, void* w, hipStream_t s) {
  for (int k = 0; k < S; ++k) {
    const size_t es = dtype == 0 ? sizeof(float) : sizeof(double);
    const char* xk = (const char*)xi + (size_t)k * n * es;
    const char* ak = xa ? (const char*)xa + (size_t)k * n * es : nullptr;
    char* wk = (char*)w + (size_t)k * 3 * n * es;
    if (dtype == 0)
      k_ps_prepare<float><<<blocks(n), PSNT, 0, s>>>(n, (const float*)xk, (const float*)ak, ps.a,
                                                     ps.q, ps.mean, ps.std, (float*)wk);
    else
      k_ps_prepare<double><<<blocks(n), PSNT, 0, s>>>(n, (const double*)xk, (const double*)ak,
                                                      ps.a, ps.q, ps.mean, ps.std, (double*)wk);
  }
  return check_hip(hipGetLastError(), "point_sources_prepare");
}

int ps_add(int dtype, int S, int E, int64_t n, const void* sky, const void* w, const void* dnu,
           void* skyt, hipStream_t s) {
  const dim3 grid(blocks(n), S);
  if (dtype == 0)
    k_ps_add<float><<<grid, PSNT, 0, s>>>(S, E, n, (const float*)sky, (const float*)w,
                                          (const float*)dnu, (float*)skyt);
  else
    k_ps_add<double><<<grid, PSNT, 0, s>>>(S, E, n, (const double*)sky, (const double*)w,
                                           (const double*)dnu, (double*)skyt);
  return check_hip(hipGetLastError(), "point_sources_add");
}

int ps_grad(int dtype, const PointSources& ps, int S, int E, int64_t n, const void* ct,
            const void* w, const void* dnu, void* g_xi, void* g_xa, hipStream_t s) {
  const dim3 grid(blocks(n), S);
  if (dtype == 0)
    k_ps_grad<float><<<grid, PSNT, 0, s>>>(S, E, n, (const float*)ct, (const float*)w,
                                           (const float*)dnu, ps.std, (float*)g_xi,
                                           (float*)g_xa);
  else
    k_ps_grad<double><<<grid, PSNT, 0, s>>>(S, E, n, (const double*)ct, (const double*)w,
                                            (const double*)dnu, ps.std, (double*)g_xi,
                                            (double*)g_xa);
  return check_hip(hipGetLastError(), "point_sources_grad");
}

}  // namespace jubik
