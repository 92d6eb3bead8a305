// Shared device-side definitions for the CReSS large-step kernels (gfx950).
//
// Layout: every 3-D field is the reference's real(0:ni+1,0:nj+1,1:nk) block
// stored exactly as Fortran does (x fastest, then y, then z), so host arrays
// move to and from the device with one contiguous copy and a wavefront of 64
// lanes reads 64 consecutive x points (256 B) per row access.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cress_hip.h"

namespace cress {

struct Grid {
  int ni, nj, nk;
  long nx;   // ni+2
  long nxy;  // (ni+2)*(nj+2)
  __host__ __device__ long at(int i, int j, int k) const {
    return (long)i + nx * (long)j + nxy * (long)(k - 1);
  }
  __host__ __device__ long n() const { return nxy * (long)nk; }
};

inline Grid make_grid(int ni, int nj, int nk) {
  Grid g;
  g.ni = ni;
  g.nj = nj;
  g.nk = nk;
  g.nx = ni + 2;
  g.nxy = (long)(ni + 2) * (nj + 2);
  return g;
}

// Scalar constants folded in binary32 exactly as the reference folds its
// REAL parameters (commath.f90, comphy.f90) and its per-call locals
// (advs.f90:189-195, advuvw.f90:209-219, stepwi.f90:308-311, ...).
struct Consts {
  float dxiv, dyiv, dziv;
  float dxv05n, dyv05n, dzv05n;       // -.5/dx ...
  float dxv25n, dyv25n, dzv25n;       // -.25/dx ...
  float dxv24, dyv24, dzv24;          // (1/24)/dx ...
  float dxv48, dyv48, dzv48;          // .5*(1/24)/dx ...
  float fourd3;
  float g05, g05n;                    // .5*g, -.5*g
  float weicoe, weic1m, filcoe, m1fc2;
  float dz;
  float thresq;
};

// Options the kernels branch on, in one POD so launches stay small.
struct Opts {
  int advopt, divopt, buyopt, bbc, tbc, sthopt, coropt, cphopt;
  int iw, ie, js, jn;  // 4th-order reduction flags (0 on periodic sides)
  int vsp_u, vsp_v, vsp_w, vsp_p, vsp_pt, vsp_qv, vsp_q;
  int ksp0;  // ksp0(2): the vspopt=2 sponge's lowest level
};

}  // namespace cress
