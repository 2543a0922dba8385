// Host-side launchers of the GP sky kernels.
#pragma once
#include "common.hpp"

namespace jubik {

struct GpGeom {
  int N;          // GP grid N x N
  int H, W;       // sky grid (crop of the GP grid at the origin)
  int E;          // energies
  int has_alpha;  // spectral-index field present
  double zm, mu;  // zero mode and spectral-index mean
  int uplanes;    // adjoint input: number of (H, W) planes summed on load (real,
                  // or complex with ucomplex)
  int S;          // samples per launch (grid axis y); inputs and outputs are
                  // stacked per sample, the amplitude fields are shared
  int sky_shared; // tangent: 1 = one base-point sky for every sample
  int work_shared; // tangent: 1 = one tangent shared by the samples (its column
                   // pass runs once; each sample's row pass reads work[0])
  int ucomplex = 0; // adjoint input: the uplanes planes are complex (H, W) planes
                    // with the energies folded in (sum_e u_e (1 + i dnu_e)),
                    // summed as they are
};

// Optional outputs of the adjoint: the gradients with respect to the
// amplitude fields (g_amp_*, N x N per sample, needing the positions xi_*)
// and to the offsets zm and mu (g_off, two values per sample).  Unused when
// g_amp_ref is NULL.
struct GpHyper {
  const void* xi_ref = nullptr;
  const void* xi_alpha = nullptr;
  void* g_amp_ref = nullptr;
  void* g_amp_alpha = nullptr;
  void* g_off = nullptr;
};

struct GpBuffers {
  const void* tw;  // N twiddles
  void* work;      // S x N x N complex
};

bool gp_length_supported(int N);
int gp_forward(int dtype, const GpGeom& g, const GpBuffers& b, const void* amp_ref,
               const void* amp_alpha, const void* xi_ref, const void* xi_alpha, const void* dnu,
               void* sky, hipStream_t s);
int gp_tangent(int dtype, const GpGeom& g, const GpBuffers& b, const void* amp_ref,
               const void* amp_alpha, const void* v_ref, const void* v_alpha, const void* dnu,
               const void* sky, void* dsky, hipStream_t s);
// Adjoint from the uplanes real (H, W) cotangent planes u of log s (per
// energy; per sample and energy when several samples are summed), plane p
// entering the spectral-index field with dnu[p mod E].
int gp_adjoint_u(int dtype, const GpGeom& g, const GpBuffers& b, const void* amp_ref,
                 const void* amp_alpha, const void* dnu, const void* u, void* g_ref,
                 void* g_alpha, hipStream_t s, GpHyper hy = GpHyper{});
// u_e = sky_e * cot_e, the E cotangent planes of log s for a cotangent of s.
int gp_fold(int dtype, const GpGeom& g, const void* sky, const void* cot, void* u,
            hipStream_t s);

}  // namespace jubik
