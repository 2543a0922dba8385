// Device-resident conjugate gradients for the MGVI / KL systems: the
// recurrences of nifty.re's CG (used by optimize_kl for the linear sample
// draws and by the Newton-CG KL minimiser, demos/x-ray-imaging.py:404-430)
// with every scalar kept in device memory, so an iteration is a fixed
// sequence of kernel launches with no host synchronisation (and can be
// captured in a graph).  Dot products are FP64 with a fixed-order
// reduction.  S independent systems run side by side; a system that meets
// its stopping rule is frozen (step length 0) while the others continue.
#pragma once
#include "common.hpp"

namespace jubik {

// A vector of the latent space: one or two fields (xi_ref, xi_alpha) of n
// elements per system, systems at a stride of n elements.
struct CgVec {
  void* f[2];
};

// Per-system state, kCgStateDoubles doubles each (see cg.hip).
constexpr int kCgStateDoubles = 12;
// Reduction blocks per system of the dot-product kernels.
constexpr int kCgBlocks = 128;

struct CgDims {
  int S;        // systems
  int nf;       // fields (1 or 2)
  int64_t n;    // elements per field and system
  double rtol;  // stop when |r| <= rtol |b|
  double absdelta;  // and/or when the energy changes by at most absdelta (<= 0: off)
};

// Workspace: state (S * kCgStateDoubles) + partials (3 * S * kCgBlocks).
int64_t cg_work_doubles(int S);

// x = 0, r = p = b, rr = bb = b.b, energy 0, active where bb > 0.
int cg_init(int dtype, const CgDims& d, CgVec b, CgVec x, CgVec r, CgVec p, double* work,
            hipStream_t s);
// Ap = c * Mv + p (c per call), pAp = p.Ap, alpha = rr / pAp (0 where frozen).
int cg_matvec(int dtype, const CgDims& d, CgVec p, CgVec Mv, double c, CgVec Ap, double* work,
              hipStream_t s);
// x += alpha p, r -= alpha Ap, rr' = r.r (and the energy -1/2 x.(b + r) for
// absdelta); stopping rule; beta = rr' / rr; p = r + beta p.
int cg_update(int dtype, const CgDims& d, CgVec b, CgVec x, CgVec r, CgVec p, CgVec Ap,
              double* work, hipStream_t s);
// info[2 s] = iterations, info[2 s + 1] = |r| / |b| (device doubles).
int cg_finish(const CgDims& d, const double* work, double* info, hipStream_t s);
// out = a + b + c elementwise (n elements; out may alias a).
int vec_sum3(int dtype, int64_t n, const void* a, const void* b, const void* c, void* out,
             hipStream_t s);
// Number of systems still iterating (synchronises the stream).
int cg_active_count(const CgDims& d, const double* work, hipStream_t s, int* count);

}  // namespace jubik
