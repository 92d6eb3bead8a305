// Open lateral boundaries on all four sides (wbc=ebc=sbc=nbc=7): radiative
// fills at a constant phase speed (rbcu, rbcv, rbcw, rbcss, rbcpt, rbcqv,
// rbcq, rbcs) and the corner averages of bc4news, for advopt<=3 and no
// lateral damping (lbcvar 'x').
//
// With option 7 every phase-speed array the reference fills in phasev
// (phvuvw.f90, phvs.f90) is one number per side over every entry the fills
// read, so the kernels take scalars:
//   acoustic u, v, w, p': c = clip(-+gwave*dtb*dxiv, +-1) * dts/dtb
//   large-step scalars:    c = clip(-+gwave*dtb*dxiv, +-1),
//                          used as gamma = 2c/(1 -+ c) (rbcpt.f90:446-470)
// where dtb, dts are the previous large step's (phasev runs at the end of a
// step) and all speeds are zero before the first step of a cold start.
#pragma once
#include "common.h"

namespace cress {

struct ObcAc {
  float cw, ce, cs, cn;   // west, east, south, north
};

struct ObcLf {
  float gw, ge, gs, gn;   // gamma per side
};

// One acoustic substep's fields: u and v are read from the buffers of the
// previous substep and written to the current ones; w and p' in place.
struct ObcAcArgs {
  const float *uold, *vold;
  float *unew, *vnew, *wf, *ppf;
};

// Up to 16 fields whose box corners get bc4news's mean of their x and y
// neighbours at every level.
struct CornerSet {
  int n;
  float *ptr[16];
  int isw[16], ise[16], jss[16], jsn[16];
};

// Large-step scalars on the boundary ring for the leapfrog without
// microphysics: theta' (no clip) and qv (clip at 0).
struct ObcStepsArgs {
  const float *s[2], *sp[2];
  float *sf[2];
};

__device__ __forceinline__ float obc_max0(bool clip0, float v) { return clip0 ? (v > 0.0f ? v : 0.0f) : v; }

// The future value of a large-step scalar at a cell of the boundary ring of
// the box 1..ni-1 x 1..nj-1 (level k, 2<=k<=nk-2): the leapfrog radiation
// of rbcpt/rbcqv/rbcq/rbcs on an edge, sf = sp -+ gamma*(s_in - sp)
// (rbcpt.f90:455-470 and its east/south/north mirrors), or at a corner the
// bc4news mean (bc4news.f90:153-204) of the two edge values next to it.
// s is the present level, sp the past one.
__device__ __forceinline__ float obc_lf_ring(const Grid &g, const ObcLf &L, bool clip0, const float *s,
                                             const float *sp, int i, int j, int k) {
  const int im = g.ni - 1, jm = g.nj - 1;
  auto W = [&](int jj) {
    const long p = g.at(1, jj, k);
    const float pv = sp[p];
    return obc_max0(clip0, pv - L.gw * (s[p + 1] - pv));
  };
  auto E = [&](int jj) {
    const long p = g.at(im, jj, k);
    const float pv = sp[p];
    return obc_max0(clip0, pv + L.ge * (s[p - 1] - pv));
  };
  auto S = [&](int ii) {
    const long p = g.at(ii, 1, k);
    const float pv = sp[p];
    return obc_max0(clip0, pv - L.gs * (s[p + g.nx] - pv));
  };
  auto N = [&](int ii) {
    const long p = g.at(ii, jm, k);
    const float pv = sp[p];
    return obc_max0(clip0, pv + L.gn * (s[p - g.nx] - pv));
  };
  if (i == 1) {
    if (j == 1) return 0.5f * (S(2) + W(2));
    if (j == jm) return 0.5f * (N(2) + W(jm - 1));
    return W(j);
  }
  if (i == im) {
    if (j == 1) return 0.5f * (S(im - 1) + E(2));
    if (j == jm) return 0.5f * (N(im - 1) + E(jm - 1));
    return E(j);
  }
  if (j == 1) return S(i);
  return N(i);
}

__global__ void k_obc_ac_edges(Grid g, ObcAc c, ObcAcArgs a);
__global__ void k_obc_corners(Grid g, CornerSet cs);
__global__ void k_obc_steps(Grid g, ObcLf lf, ObcStepsArgs a);

}  // namespace cress
