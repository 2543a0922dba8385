// eROSITA PSF interpolation onto the linear-patching patches
// (jubik/instruments/erosita/erosita_psf_utils.py:90-264): each patch PSF is
// the tabulated PSFs at the patch's off-axis radius (linear weights over the
// radii), rotated by the patch's azimuth and sampled bilinearly with the
// index clamping of map_coordinates(order=1, mode='nearest').
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "psf.hpp"

namespace jubik {

constexpr int PNT = 256;

// grid (ceil(npix / 256), P).  layout 0: full frame (H, W) rolled by
// (-H/2, -W/2); layout 1: the centred w x w window linpatch cuts from it.
template <typename T>
__global__ __launch_bounds__(PNT) void k_psf_interp(PsfGeom g, const T* __restrict__ psfs,
                                                    const int* __restrict__ order,
                                                    const double* __restrict__ cid,
                                                    const double* __restrict__ phi,
                                                    const double* __restrict__ wgt,
                                                    T* __restrict__ out) {
  const int p = blockIdx.y;
  const int64_t pix = (int64_t)blockIdx.x * PNT + threadIdx.x;
  const int64_t npix = g.layout == 0 ? (int64_t)g.H * g.W : (int64_t)g.w * g.w;
  if (pix >= npix) return;
  int i, j;
  int64_t o;
  if (g.layout == 0) {
    i = (int)(pix / g.W);
    j = (int)(pix - (int64_t)i * g.W);
    const int ri = (i - g.H / 2 + g.H) % g.H, rj = (j - g.W / 2 + g.W) % g.W;
    o = (((int64_t)p * g.E + g.e) * g.H + ri) * g.W + rj;
  } else {
    const int r = (int)(pix / g.w), c = (int)(pix - (int64_t)r * g.w);
    i = (g.H - g.w) / 2 + r;
    j = (g.W - g.w) / 2 + c;
    o = (((int64_t)p * g.E + g.e) * g.w + r) * g.w + c;
  }
  // offsets of the pixel from the frame centre (d_dec along rows, d_ra
  // against columns), their polar form, the rotation by -phi_patch, back
  // to Cartesian (to_r_phi / to_ra_dec)
  const double d_dec = (double)(i - g.H / 2) * g.dH;
  const double d_ra = -((double)(j - g.W / 2) * g.dW);
  const double kHalfPi = 1.57079632679489661923;
  const double rr = sqrt(d_ra * d_ra + d_dec * d_dec);
  const double ang = atan2(d_dec, d_ra) - kHalfPi - phi[p];
  const double x = rr * cos(ang + kHalfPi);
  const double y = rr * sin(ang + kHalfPi);
  double acc = 0.0;
  for (int k = 0; k < g.nrad; ++k) {
    const double wk = wgt[(int64_t)p * g.nrad + k];
    if (wk == 0.0) continue;
    // to_patch_coordinates: (offset + centre * delta) / delta
    const double q0 = (x + cid[2 * k] * g.d0) / g.d0;
    const double q1 = (y + cid[2 * k + 1] * g.d1) / g.d1;
    const double f0 = floor(q0), f1 = floor(q1);
    const double t0 = q0 - f0, t1 = q1 - f1;
    const int a0 = (int)fmin(fmax(f0, 0.0), (double)(g.ph - 1));
    const int a1 = (int)fmin(fmax(f0 + 1.0, 0.0), (double)(g.ph - 1));
    const int b0 = (int)fmin(fmax(f1, 0.0), (double)(g.pw - 1));
    const int b1 = (int)fmin(fmax(f1 + 1.0, 0.0), (double)(g.pw - 1));
    const T* P = psfs + (int64_t)order[k] * g.ph * g.pw;
    const double v = (1.0 - t0) * ((1.0 - t1) * (double)P[(int64_t)a0 * g.pw + b0] +
                                   t1 * (double)P[(int64_t)a0 * g.pw + b1]) +
                     t0 * ((1.0 - t1) * (double)P[(int64_t)a1 * g.pw + b0] +
                           t1 * (double)P[(int64_t)a1 * g.pw + b1]);
    acc += wk * v;
  }
  out[o] = (T)acc;
}

int psf_interpolate(int dtype, const PsfGeom& g, const void* psfs, const int* order,
                    const double* cid, const double* phi, const double* wgt, void* out,
                    hipStream_t s) {
  const int64_t npix = g.layout == 0 ? (int64_t)g.H * g.W : (int64_t)g.w * g.w;
  dim3 grid((unsigned)((npix + PNT - 1) / PNT), (unsigned)(g.npatch * g.npatch));
  if (dtype == 0)
    k_psf_interp<float><<<grid, PNT, 0, s>>>(g, (const float*)psfs, order, cid, phi, wgt,
                                              (float*)out);
  else
    k_psf_interp<double><<<grid, PNT, 0, s>>>(g, (const double*)psfs, order, cid, phi, wgt,
                                               (double*)out);
  return check_hip(hipGetLastError(), "psf_interpolate");
}

// Host set-up of one energy's patch table (double, following the
// reference): sort the radii, patch centres from the meshgrid ('xy',
// flipped ra), their polar coordinates about the pointing centre and the
// radial weights (get_interpolation_weights, the three cases summed).
void psf_patch_table(const double* rs, int nrad, const double* pointing_center, int H, int W,
                     double dH, double dW, int npatch, std::vector<int>& order,
                     std::vector<double>& phi, std::vector<double>& wgt) {
  order.resize(nrad);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rs[a] < rs[b]; });
  std::vector<double> r(nrad);
  for (int k = 0; k < nrad; ++k) r[k] = rs[order[k]];
  const int P = npatch * npatch;
  const double ps0 = H / npatch, ps1 = W / npatch;
  phi.assign(P, 0.0);
  wgt.assign((size_t)P * nrad, 0.0);
  const double kHalfPi = 1.57079632679489661923;
  for (int p = 0; p < P; ++p) {
    const double dec = ((p % npatch) * ps0 + ps0 / 2) * dH;
    const double ra = (((P - 1 - p) / npatch) * ps1 + ps1 / 2) * dW;
    const double cx = ra - pointing_center[0], cy = dec - pointing_center[1];
    const double rp = std::sqrt(cx * cx + cy * cy);
    phi[p] = std::atan2(cy, cx) - kHalfPi;
    double* w = &wgt[(size_t)p * nrad];
    if (rp <= r[0]) w[0] += 1.0;
    if (rp >= r[nrad - 1]) w[nrad - 1] += 1.0;
    for (int k = 0; k + 1 < nrad; ++k)
      if (r[k] < rp && r[k + 1] >= rp) {
        const double t = (rp - r[k]) / (r[k + 1] - r[k]);
        w[k] += 1.0 - t;
        w[k + 1] += t;
      }
  }
}

}  // namespace jubik
