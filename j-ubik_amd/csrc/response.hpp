// Host-side launchers of the exposure, read-out and Poisson kernels.
#pragma once
#include "common.hpp"

namespace jubik {

constexpr int64_t kPoissonWorkDoubles = 1024;

int exposure_apply(int dtype, const void* expo, int64_t n_e, const void* x, int64_t n_x,
                   void* out, int64_t n_out, hipStream_t s);
int readout_gather(int dtype, const void* x, const int64_t* idx, int64_t nnz, void* out,
                   hipStream_t s);
int readout_scatter(int dtype, const void* v, const int64_t* idx, int64_t nnz, int64_t n,
                    void* out, hipStream_t s);
// counts[i] ~ Poisson(lam[i]), a pure function of (seed, i, lam[i]).
int poisson_sample(int dtype, const void* lam, int64_t n, uint64_t seed, int* counts,
                   hipStream_t s);
// cot[s, i] = expo[i] noise[s, i] / sqrt(lam[i]) where counts[i] >= 0, else 0
// (n pixels per sample, S samples).
int sqrt_metric_cot(int dtype, const void* expo, const int* counts, const void* lam,
                    const void* noise, void* cot, int64_t n, int S, hipStream_t s);
// geoVI: geo = sqrt(lam_x lam_p), eta = 2 (sqrt lam_x - sqrt lam_p) where
// counts >= 0 (else geo = 1, eta = 0); n pixels.
int geo_prep(int dtype, const int* counts, const void* lam_p, const void* lam_x, void* geo,
             void* eta, int64_t n, hipStream_t s);
// out = t / lam (sqrt_form 0) or t / sqrt(lam) (sqrt_form 1).
int poisson_metric(int dtype, const void* lam, const void* t, int64_t n, int sqrt_form,
                   void* out, hipStream_t s);
int poisson_nll_grad(int dtype, const void* lam, const void* d, int64_t n, double* nll,
                     void* grad, double* work, hipStream_t s);
// poisson_nll_grad with the partials in the library's static device scratch
// (no allocation; calls that overlap on one GPU race on it).
int poisson_nll_grad_noalloc(int dtype, const void* lam, const void* d, int64_t n, double* nll,
                             void* grad, hipStream_t s);

}  // namespace jubik
