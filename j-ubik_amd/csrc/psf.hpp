// Host launcher of the eROSITA PSF interpolation kernel.
#pragma once
#include <vector>

#include "common.hpp"

namespace jubik {

struct PsfGeom {
  int H, W;        // sky grid
  double dH, dW;   // pixel sizes (domain distances)
  int npatch;      // patches per axis (P = npatch^2)
  int nrad;        // tabulated off-axis radii
  int ph, pw;      // shape of each tabulated PSF
  double d0, d1;   // PSF pixel sizes (patch_deltas)
  int layout;      // 0: full frame rolled to the origin; 1: centred w x w window
  int w;           // window width (layout 1)
  int e, E;        // energy slot written and number of energies
};

int psf_interpolate(int dtype, const PsfGeom& g, const void* psfs, const int* order,
                    const double* cid, const double* phi, const double* wgt, void* out,
                    hipStream_t s);
void psf_patch_table(const double* rs, int nrad, const double* pointing_center, int H, int W,
                     double dH, double dW, int npatch, std::vector<int>& order,
                     std::vector<double>& phi, std::vector<double>& wgt);

}  // namespace jubik
