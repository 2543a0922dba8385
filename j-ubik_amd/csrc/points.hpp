// Point-source sky component (points.hip): inverse-gamma reference map and a
// per-pixel spectral index.
#pragma once
#include "common.hpp"

namespace jubik {

struct PointSources {
  double a = 0.0, q = 0.0;        // inverse-gamma shape and scale
  double mean = 0.0, std = 1.0;   // normal prior of the spectral index
};

// logi[i] = log I(xi[i]) (I = invgamma_prior(a, q)), dlogi[i] = d log I / d xi
// (dlogi may be NULL); n pixels.
int invgamma_prior(int dtype, const void* xi, int64_t n, double a, double q, void* logi,
                   void* dlogi, hipStream_t s);
// w (S, 3, n): log I, d log I / d xi and alpha = mean + std xi_alpha (0 when
// xa is NULL) of the S positions xi, xa (S, n).
int ps_prepare(int dtype, const PointSources& ps, int S, int64_t n, const void* xi,
               const void* xa, void* w, hipStream_t s);
// skyt (S, E, n) = sky + exp(log I + alpha dnu_e) (dnu NULL: no spectral part).
int ps_add(int dtype, int S, int E, int64_t n, const void* sky, const void* w, const void* dnu,
           void* skyt, hipStream_t s);
// Gradients with respect to xi and xi_alpha (S, n) from the cotangent ct
// (S, E, n) of the total sky; g_xa may be NULL.
int ps_grad(int dtype, const PointSources& ps, int S, int E, int64_t n, const void* ct,
            const void* w, const void* dnu, void* g_xi, void* g_xa, hipStream_t s);

}  // namespace jubik
