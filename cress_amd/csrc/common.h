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
  // Base-state and metric fields (ids CRESS_JCB..CRESS_RCSQ) whose uploaded
  // columns are all equal -- flat terrain and a horizontally uniform base
  // state -- are also held as one value per level: tab[id*nk + k-1], bit id
  // of umask set.  Kernels read those through MP/MT/MQ, which take the
  // table (a wave-uniform address, served by the scalar cache) instead of
  // streaming the 3-D field; the values are the same bits either way.
  const float *tab;
  unsigned umask;
  __host__ __device__ long at(int i, int j, int k) const {
    return (long)i + nx * (long)j + nxy * (long)(k - 1);
  }
  __host__ __device__ long n() const { return nxy * (long)nk; }
  // byte offset of (i, j, k) in 32 bits (fields are below 4 GiB): the
  // scalar-base + 32-bit vector-offset addressing form, one VGPR per
  // address and 32-bit integer arithmetic for the neighbours' offsets
  __host__ __device__ uint32_t off(int i, int j, int k) const {
    return 4u * ((uint32_t)i + (uint32_t)nx * (uint32_t)j + (uint32_t)nxy * (uint32_t)(k - 1));
  }
};

// Element at a 32-bit byte offset from a field's base.
__host__ __device__ __forceinline__ const float &at32(const float *a, uint32_t o) {
  return *(const float *)((const char *)a + o);
}
__host__ __device__ __forceinline__ float &at32(float *a, uint32_t o) { return *(float *)((char *)a + o); }

inline Grid make_grid(int ni, int nj, int nk) {
  Grid g;
  g.ni = ni;
  g.nj = nj;
  g.nk = nk;
  g.nx = ni + 2;
  g.nxy = (long)(ni + 2) * (nj + 2);
  g.tab = nullptr;
  g.umask = 0u;
  return g;
}

// The device-resident context's fields: the same (i, j, k) order, rows
// padded to a multiple of CRESS_ROW_ALIGN floats, so rows start on aligned
// lines and the kernels' 64-point x tiles from i = 64t do not straddle an
// extra line.  Host arrays and the per-routine entry points keep the
// unpadded (ni+2) rows.  Measured at 1024x1024x80 (ms per step, 6 steps):
// unpadded 43.8, 16-float rows 43.2, 32-float rows 43.3 (32 helps the
// scalar forcing most, 3.87 -> 3.44 ms, but slows the Asselin/saturation
// kernel, 4.03 -> 4.34 ms; 16 sits between on both).
#ifndef CRESS_ROW_ALIGN
#define CRESS_ROW_ALIGN 16
#endif
#ifndef CRESS_ROW_EXTRA
#define CRESS_ROW_EXTRA 0
#endif
inline Grid make_grid_padded(int ni, int nj, int nk) {
  Grid g = make_grid(ni, nj, nk);
  g.nx = (long)((ni + 2 + CRESS_ROW_ALIGN - 1) / CRESS_ROW_ALIGN * CRESS_ROW_ALIGN + CRESS_ROW_EXTRA);
  g.nxy = g.nx * (long)(nj + 2);
  return g;
}

// A base-state / metric field at linear index p (MP) or (i,j) (MT) on level
// k; MQ: at byte offset q, for the kernels that address by offsets.  Need a
// Grid named g in scope.
#define MUNI(id) ((g.umask >> (id)) & 1u)
#define MTAB(id, k) g.tab[(id) * g.nk + (k) - 1]
#define MP(id, arr, p, k) (MUNI(id) ? MTAB(id, k) : (arr)[p])
#define MT(id, arr, i, j, k) (MUNI(id) ? MTAB(id, k) : at32(arr, g.off(i, j, k)))
#define MQ(id, arr, q, k) (MUNI(id) ? MTAB(id, k) : *(const float *)((const char *)(arr) + (q)))
// In a kernel templated on U (true: every base/metric field it reads is a
// per-level table), the table read needs no run-time test, so the compiler
// keeps it a scalar load instead of selecting between two addresses.
#define MQU(id, arr, q, k) (U ? MTAB(id, k) : MQ(id, arr, q, k))
#define MTU(id, arr, i, j, k) (U ? MTAB(id, k) : MT(id, arr, i, j, k))
#define MPU(id, arr, p, k) (U ? MTAB(id, k) : MP(id, arr, p, k))

// Flattened launches with an XCD-aware tile order.  The workgroups of a
// launch go to the 8 XCDs round-robin (workgroup b to XCD b % 8), and each
// XCD has its own L2.  xcd_tile() maps workgroup b to a logical tile so that
// every XCD walks one contiguous eighth of the (x fastest, then y, then z)
// tile space: stencil neighbours then meet in the same L2 instead of being
// fetched once per XCD.  Order 2 walks z fastest inside each XCD's chunk,
// so the k-neighbours of a 3-D stencil are the next tiles on the same L2.
struct Tiles {
  unsigned x, y, z;
  unsigned order;  // 0: plain (no remap), 1: XCD chunks x-fastest, 2: XCD chunks z-fastest
};

// Order per kernel, measured at 1024x1024x80: 3-D stencils with k
// neighbours (forcing, ptil) run best z-fastest, pointwise and column
// kernels x-fastest.
inline Tiles tiles(int nx, int ny, int bx, int by, int nz, unsigned order = 1) {
  Tiles t;
  t.x = (unsigned)((nx + bx - 1) / bx);
  t.y = (unsigned)((ny + by - 1) / by);
  t.z = (unsigned)nz;
  t.order = order;
  return t;
}

inline dim3 flat(const Tiles &t) { return dim3(t.x * t.y * t.z); }

__device__ __forceinline__ uint3 xcd_tile(const Tiles t) {
  const unsigned nb = t.x * t.y * t.z;
  const unsigned b = blockIdx.x;
  unsigned l = b;
  if (t.order != 0) {
    const unsigned x = b & 7u, r = b >> 3;
    const unsigned q = nb >> 3, rem = nb & 7u;
    l = x * q + (x < rem ? x : rem) + r;
  }
  uint3 o;
  if (t.order == 2) {
    o.z = l % t.z;
    const unsigned s = l / t.z;
    o.x = s % t.x;
    o.y = s / t.x;
  } else {
    o.x = l % t.x;
    const unsigned s = l / t.x;
    o.y = s % t.y;
    o.z = s / t.y;
  }
  return o;
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
  float smh, smv;                     // smoothing coefficients smhcoe, smvcoe
};

// Options the kernels branch on, in one POD so launches stay small.
struct Opts {
  int advopt, divopt, buyopt, bbc, tbc, sthopt, coropt, cphopt;
  int iw, ie, js, jn;  // 4th-order reduction flags (0 on periodic sides)
  int vsp_u, vsp_v, vsp_w, vsp_p, vsp_pt, vsp_qv, vsp_q;
  int ksp0;  // ksp0(2): the vspopt=2 sponge's lowest level
  int smtopt;  // 0, or 2/3: 4th-order smoothing horizontal / also vertical
};

}  // namespace cress
