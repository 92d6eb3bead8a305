// Dynamics kernels of one CReSS large step for gfx950: forcing (advection,
// sponge, Coriolis, buoyancy), the HE-VI acoustic substep, the scalar
// leapfrog update, the Asselin filter and the on-device halo fills.
//
// Every kernel restates a reference routine's arithmetic with the same
// operation order (the build uses -ffp-contract=off and IEEE division), so
// the FP32 results equal the reference's FP32 results bitwise.  Ranges are
// the reference's loop ranges; the routine each kernel replaces is cited.
#include "common.h"
#include "dyn.h"

namespace cress {

#define AT(a, i, j, k) at32(a, g.off(i, j, k))
// Load / store at a 32-bit byte offset from a kernel-argument base: the
// scalar-base + vector-offset addressing form, one VGPR per address.
#define LD(arr, off) (*(const float *)((const char *)(arr) + (off)))
#define ST(arr, off, v) (*(float *)((char *)(arr) + (off)) = (v))

// rbct(1:ni,1:nj,1:nk,1:2), component 2 (vspdmp.f90:269-285)
__device__ __forceinline__ float rbct2(const float *rbct, const Grid &g, int i, int j, int k) {
  return at32(rbct, 4u * ((uint32_t)(i - 1) + (uint32_t)g.ni * ((uint32_t)(j - 1) + (uint32_t)g.nj *
                                                                 ((uint32_t)(k - 1) + (uint32_t)g.nk))));
}

// vspuvw.f90:222-238: sponge coefficient averaged to the u/v level
__device__ __forceinline__ float rbct8s(const float *rbct, const Grid &g, int i, int j, int k) {
  return 0.25f * (rbct2(rbct, g, i, j, k) + rbct2(rbct, g, i, j, k + 1));
}

// 4th-order numerical smoothing of one output point (smoo4s.f90:187-298 and
// its u/v/w, qv and p' twins in smoo4uvw.f90, smoo4qv.f90, smoo4p.f90):
// a(di, dj, dk) is the prepared field at an offset (rbr*s, rst8u*(u-ubr)/
// jcb8u, ...).  The 2nd-order part f + smhcoe*(t1+t2) + smvcoe*t3 applies
// everywhere; the 4th-order part, built from the neighbours' second
// differences, inside the band where the 4th-order advection applies.
// kb / kt: the point is the lowest / highest level of its range, where the
// reference continues t3 by copy (smoo4s.f90:270-275).
template <class A>
__device__ __forceinline__ float smooth4(float f, const Consts &c, int smtopt, bool band, bool kb, bool kt,
                                         A a) {
  const float a0 = a(0, 0, 0);
  const float t1 = (a(1, 0, 0) + a(-1, 0, 0)) - 2.0f * a0;
  const float t2 = (a(0, 1, 0) + a(0, -1, 0)) - 2.0f * a0;
  const float t3 = (a(0, 0, 1) + a(0, 0, -1)) - 2.0f * a0;
  f = f + c.smh * (t1 + t2) + c.smv * t3;
  if (!band) return f;
  auto T1 = [&](int d) { return (a(d + 1, 0, 0) + a(d - 1, 0, 0)) - 2.0f * a(d, 0, 0); };
  auto T2 = [&](int d) { return (a(0, d + 1, 0) + a(0, d - 1, 0)) - 2.0f * a(0, d, 0); };
  auto T3 = [&](int d) { return (a(0, 0, d + 1) + a(0, 0, d - 1)) - 2.0f * a(0, 0, d); };
  const float h = (t1 - (T1(1) + T1(-1))) + (t2 - (T2(1) + T2(-1)));
  if (smtopt == 2) return f + c.smh * h;
  const float t3p = kt ? t3 : T3(1), t3m = kb ? t3 : T3(-1);
  return f + (c.smv * (t3 - (t3p + t3m)) + c.smh * h);
}

// ---------------------------------------------------------------------------
// Contravariant w (phy2cnt.f90:193-227 then vbcwc.f90:130-186) and the
// mass-flux velocities rstx* (rstuvwc.f90:172-210).
// ---------------------------------------------------------------------------
template <bool U>
__device__ __forceinline__ float wc_raw(const Grid &g, int sthopt, const float *w,
                                        const float *jcb8w, int i, int j, int k) {
  return sthopt == 0 ? AT(w, i, j, k) : AT(w, i, j, k) / MTU(CRESS_JCB8W, jcb8w, i, j, k);
}

template <bool U>
__device__ __forceinline__ float wc_value(const Grid &g, const Opts &o, const float *w,
                                          const float *jcb8w, int i, int j, int k) {
  const int nk = g.nk;
  if (k == 1) {
    if (o.bbc == 2) return -wc_raw<U>(g, o.sthopt, w, jcb8w, i, j, 3);
    return wc_raw<U>(g, o.sthopt, w, jcb8w, i, j, 2);
  }
  if (k == 2 && o.bbc == 2) return 0.0f;
  if (k == nk) {
    if (o.tbc == 2) return -wc_raw<U>(g, o.sthopt, w, jcb8w, i, j, nk - 2);
    return wc_raw<U>(g, o.sthopt, w, jcb8w, i, j, nk - 1);
  }
  if (k == nk - 1 && o.tbc == 2) return 0.0f;
  return wc_raw<U>(g, o.sthopt, w, jcb8w, i, j, k);
}

// wc over i=1..ni-1, j=1..nj-1, k=1..nk (the phy2cnt+vbcwc write range)
template <bool U>
__global__ void k_phy2cnt(Tiles T, Grid g, Opts o, const float *w, const float *jcb8w, float *wc) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 1 + xt.y * blockDim.y + threadIdx.y;
  const int k = 1 + xt.z;
  if (i < 1 || i > g.ni - 1 || j > g.nj - 1) return;
  AT(wc, i, j, k) = wc_value<U>(g, o, w, jcb8w, i, j, k);
}

template <bool U>
__global__ void k_prep(Tiles T, Grid g, Opts o, const float *u, const float *v, const float *w,
                       const float *rst8u, const float *rst8v, const float *rst8w,
                       const float *jcb8w, float *wc, float *rstxu, float *rstxv, float *rstxwc) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = xt.y * blockDim.y + threadIdx.y;
  const int k = 1 + xt.z;
  if (i > g.ni + 1 || j > g.nj + 1) return;
  // every value is formed before the first store, so no load waits on one
  const bool dw = i >= 1 && i <= g.ni - 1 && j >= 1 && j <= g.nj - 1;
  const bool du = k <= g.nk - 1 && j >= 1 && j <= g.nj - 1 && i >= o.iw && i <= g.ni + 1 - o.ie;
  const bool dv = k <= g.nk - 1 && i >= 1 && i <= g.ni - 1 && j >= o.js && j <= g.nj + 1 - o.jn;
  float c = 0.0f, xw = 0.0f, xu = 0.0f, xv = 0.0f;
  if (dw) {
    c = wc_value<U>(g, o, w, jcb8w, i, j, k);
    xw = MTU(CRESS_RST8W, rst8w, i, j, k) * c;
  }
  if (du) xu = MTU(CRESS_RST8U, rst8u, i, j, k) * AT(u, i, j, k);
  if (dv) xv = MTU(CRESS_RST8V, rst8v, i, j, k) * AT(v, i, j, k);
  if (dw) {
    AT(wc, i, j, k) = c;
    AT(rstxwc, i, j, k) = xw;
  }
  if (du) AT(rstxu, i, j, k) = xu;
  if (dv) AT(rstxv, i, j, k) = xv;
}

// ---------------------------------------------------------------------------
// Momentum advection, advuvw.f90:209-902.  Each returns f + advection with
// the reference's accumulation order: advopt=1 f+(...), advopt=2
// (f+hadv)+(t3+t3), advopt=3 (f+hadv)+vadv.
// ---------------------------------------------------------------------------
__device__ float adv_u(float f, const Grid &g, const Opts &o, const Consts &c, const float *u,
                       const float *rxu, const float *rxv, const float *rxw, int i, int j, int k) {
  auto t1 = [&](int ii) {
    return (AT(rxu, ii, j, k) + AT(rxu, ii + 1, j, k)) * (AT(u, ii + 1, j, k) - AT(u, ii, j, k)) * c.dxv25n;
  };
  auto t2 = [&](int jj) {
    return (AT(rxv, i - 1, jj, k) + AT(rxv, i, jj, k)) * (AT(u, i, jj, k) - AT(u, i, jj - 1, k)) * c.dyv25n;
  };
  auto t3 = [&](int kk) {
    return (AT(rxw, i - 1, j, kk) + AT(rxw, i, j, kk)) * (AT(u, i, j, kk) - AT(u, i, j, kk - 1)) * c.dzv25n;
  };
  if (o.advopt == 1) return f + ((t3(k) + t3(k + 1)) + ((t1(i - 1) + t1(i)) + (t2(j) + t2(j + 1))));
  float hadv = (t1(i - 1) + t1(i)) + (t2(j) + t2(j + 1));
  const bool h4 = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 1 - o.ie;
  if (h4) {
    auto T1 = [&](int ii) {
      return (AT(rxu, ii - 1, j, k) + AT(rxu, ii + 1, j, k)) * (AT(u, ii + 1, j, k) - AT(u, ii - 1, j, k)) * c.dxv24;
    };
    auto T2 = [&](int jj) {
      return ((AT(rxv, i - 1, jj, k) + AT(rxv, i, jj, k)) + (AT(rxv, i - 1, jj + 1, k) + AT(rxv, i, jj + 1, k))) *
             (AT(u, i, jj + 1, k) - AT(u, i, jj - 1, k)) * c.dyv48;
    };
    hadv = c.fourd3 * hadv + ((T1(i - 1) + T1(i + 1)) + (T2(j - 1) + T2(j + 1)));
  }
  if (o.advopt == 2) return f + hadv + (t3(k) + t3(k + 1));
  float vadv = t3(k) + t3(k + 1);
  if (h4 && k >= 3 && k <= g.nk - 3) {
    auto T3 = [&](int kk) {
      return ((AT(rxw, i - 1, j, kk) + AT(rxw, i, j, kk)) + (AT(rxw, i - 1, j, kk + 1) + AT(rxw, i, j, kk + 1))) *
             (AT(u, i, j, kk + 1) - AT(u, i, j, kk - 1)) * c.dzv48;
    };
    vadv = c.fourd3 * vadv + (T3(k - 1) + T3(k + 1));
  }
  return f + hadv + vadv;
}

__device__ float adv_v(float f, const Grid &g, const Opts &o, const Consts &c, const float *v,
                       const float *rxu, const float *rxv, const float *rxw, int i, int j, int k) {
  auto t1 = [&](int ii) {
    return (AT(rxu, ii, j - 1, k) + AT(rxu, ii, j, k)) * (AT(v, ii, j, k) - AT(v, ii - 1, j, k)) * c.dxv25n;
  };
  auto t2 = [&](int jj) {
    return (AT(rxv, i, jj, k) + AT(rxv, i, jj + 1, k)) * (AT(v, i, jj + 1, k) - AT(v, i, jj, k)) * c.dyv25n;
  };
  auto t3 = [&](int kk) {
    return (AT(rxw, i, j - 1, kk) + AT(rxw, i, j, kk)) * (AT(v, i, j, kk) - AT(v, i, j, kk - 1)) * c.dzv25n;
  };
  if (o.advopt == 1) return f + ((t3(k) + t3(k + 1)) + ((t1(i) + t1(i + 1)) + (t2(j - 1) + t2(j))));
  float hadv = (t1(i) + t1(i + 1)) + (t2(j - 1) + t2(j));
  const bool h4 = j >= 2 + o.js && j <= g.nj - 1 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
  if (h4) {
    auto T1 = [&](int ii) {
      return ((AT(rxu, ii, j - 1, k) + AT(rxu, ii, j, k)) + (AT(rxu, ii + 1, j - 1, k) + AT(rxu, ii + 1, j, k))) *
             (AT(v, ii + 1, j, k) - AT(v, ii - 1, j, k)) * c.dxv48;
    };
    auto T2 = [&](int jj) {
      return (AT(rxv, i, jj - 1, k) + AT(rxv, i, jj + 1, k)) * (AT(v, i, jj + 1, k) - AT(v, i, jj - 1, k)) * c.dyv24;
    };
    hadv = c.fourd3 * hadv + ((T1(i - 1) + T1(i + 1)) + (T2(j - 1) + T2(j + 1)));
  }
  if (o.advopt == 2) return f + hadv + (t3(k) + t3(k + 1));
  float vadv = t3(k) + t3(k + 1);
  if (h4 && k >= 3 && k <= g.nk - 3) {
    auto T3 = [&](int kk) {
      return ((AT(rxw, i, j - 1, kk) + AT(rxw, i, j, kk)) + (AT(rxw, i, j - 1, kk + 1) + AT(rxw, i, j, kk + 1))) *
             (AT(v, i, j, kk + 1) - AT(v, i, j, kk - 1)) * c.dzv48;
    };
    vadv = c.fourd3 * vadv + (T3(k - 1) + T3(k + 1));
  }
  return f + hadv + vadv;
}

__device__ float adv_w(float f, const Grid &g, const Opts &o, const Consts &c, const float *w,
                       const float *rxu, const float *rxv, const float *rxw, int i, int j, int k) {
  auto t1 = [&](int ii) {
    return (AT(rxu, ii, j, k - 1) + AT(rxu, ii, j, k)) * (AT(w, ii, j, k) - AT(w, ii - 1, j, k)) * c.dxv25n;
  };
  auto t2 = [&](int jj) {
    return (AT(rxv, i, jj, k - 1) + AT(rxv, i, jj, k)) * (AT(w, i, jj, k) - AT(w, i, jj - 1, k)) * c.dyv25n;
  };
  auto t3 = [&](int kk) {
    return (AT(rxw, i, j, kk) + AT(rxw, i, j, kk + 1)) * (AT(w, i, j, kk + 1) - AT(w, i, j, kk)) * c.dzv25n;
  };
  if (o.advopt == 1) return f + ((t3(k - 1) + t3(k)) + ((t1(i) + t1(i + 1)) + (t2(j) + t2(j + 1))));
  float hadv = (t1(i) + t1(i + 1)) + (t2(j) + t2(j + 1));
  const bool h4 = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
  if (h4) {
    auto T1 = [&](int ii) {
      return ((AT(rxu, ii, j, k - 1) + AT(rxu, ii, j, k)) + (AT(rxu, ii + 1, j, k - 1) + AT(rxu, ii + 1, j, k))) *
             (AT(w, ii + 1, j, k) - AT(w, ii - 1, j, k)) * c.dxv48;
    };
    auto T2 = [&](int jj) {
      return ((AT(rxv, i, jj, k - 1) + AT(rxv, i, jj, k)) + (AT(rxv, i, jj + 1, k - 1) + AT(rxv, i, jj + 1, k))) *
             (AT(w, i, jj + 1, k) - AT(w, i, jj - 1, k)) * c.dyv48;
    };
    hadv = c.fourd3 * hadv + ((T1(i - 1) + T1(i + 1)) + (T2(j - 1) + T2(j + 1)));
  }
  if (o.advopt == 2) return f + hadv + (t3(k - 1) + t3(k));
  float vadv = t3(k - 1) + t3(k);
  if (h4 && k >= 3 && k <= g.nk - 2) {
    auto T3 = [&](int kk) {
      return (AT(rxw, i, j, kk - 1) + AT(rxw, i, j, kk + 1)) * (AT(w, i, j, kk + 1) - AT(w, i, j, kk - 1)) * c.dzv24;
    };
    vadv = c.fourd3 * vadv + (T3(k - 1) + T3(k + 1));
  }
  return f + hadv + vadv;
}

// ufrc over i=2..ni-1, j=2..nj-2 and vfrc over i=2..ni-2, j=2..nj-1; k=2..nk-2.
// Accumulation order of forceuvw.f90:434-626: 0, smoothing, sponge,
// Coriolis, advection.
template <bool SM, bool U>
__global__ void k_force_uv(Tiles T, Grid g, Opts o, Consts c, ForceIn in, float *ufrc, float *vfrc) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 2 + xt.y * blockDim.y + threadIdx.y;
  const int k = 2 + xt.z;
  if (i < 2 || i > g.ni - 1 || j > g.nj - 1) return;
  const bool sp = k >= o.ksp0 - 1 && k <= g.nk - 2;
  // both forcings are formed before either is stored, so that v's loads
  // need not wait for u's store
  const bool du = j <= g.nj - 2, dv = i <= g.ni - 2;
  float uout = 0.0f, vout = 0.0f;
  if (du) {
    float f = 0.0f;
    if (SM) {
      const bool band = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 1 - o.ie;
      f = smooth4(f, c, o.smtopt, band, k == 2, k == g.nk - 2, [&](int di, int dj, int dk) {
        const long p = g.at(i + di, j + dj, k + dk);
        const int kk = k + dk;
        return MPU(CRESS_RST8U, in.rst8u, p, kk) * (in.up[p] - MPU(CRESS_UBR, in.ubr, p, kk)) /
               MPU(CRESS_JCB8U, in.jcb8u, p, kk);
      });
    }
    if (o.vsp_u && sp)
      f = f - MTU(CRESS_RST8U, in.rst8u, i, j, k) * (rbct8s(in.rbct, g, i - 1, j, k) + rbct8s(in.rbct, g, i, j, k)) *
                  (AT(in.up, i, j, k) - MTU(CRESS_UBR, in.ubr, i, j, k));
    if (o.coropt) {
      auto cr = [&](int ii) {
        return at32(in.fc, 4u * ((uint32_t)ii + (uint32_t)(g.ni + 2) * (uint32_t)j)) * MTU(CRESS_RST, in.rst, ii, j, k) *
               (AT(in.v, ii, j, k) + AT(in.v, ii, j + 1, k));
      };
      f = f + (cr(i - 1) + cr(i));
    }
    uout = adv_u(f, g, o, c, in.u, in.rstxu, in.rstxv, in.rstxwc, i, j, k);
  }
  if (dv) {
    float f = 0.0f;
    if (SM) {
      const bool band = j >= 2 + o.js && j <= g.nj - 1 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
      f = smooth4(f, c, o.smtopt, band, k == 2, k == g.nk - 2, [&](int di, int dj, int dk) {
        const long p = g.at(i + di, j + dj, k + dk);
        const int kk = k + dk;
        return MPU(CRESS_RST8V, in.rst8v, p, kk) * (in.vp[p] - MPU(CRESS_VBR, in.vbr, p, kk)) /
               MPU(CRESS_JCB8V, in.jcb8v, p, kk);
      });
    }
    if (o.vsp_v && sp)
      f = f - MTU(CRESS_RST8V, in.rst8v, i, j, k) * (rbct8s(in.rbct, g, i, j - 1, k) + rbct8s(in.rbct, g, i, j, k)) *
                  (AT(in.vp, i, j, k) - MTU(CRESS_VBR, in.vbr, i, j, k));
    if (o.coropt) {
      auto cr = [&](int jj) {
        return -at32(in.fc, 4u * ((uint32_t)i + (uint32_t)(g.ni + 2) * (uint32_t)jj)) * MTU(CRESS_RST, in.rst, i, jj, k) *
               (AT(in.u, i, jj, k) + AT(in.u, i + 1, jj, k));
      };
      f = f + (cr(j - 1) + cr(j));
    }
    vout = adv_v(f, g, o, c, in.v, in.rstxu, in.rstxv, in.rstxwc, i, j, k);
  }
  if (du) AT(ufrc, i, j, k) = uout;
  if (dv) AT(vfrc, i, j, k) = vout;
}

// moist buoyancy at a scalar level, buoywb.f90:204-297 (gwmopt=0)
template <bool U>
__device__ __forceinline__ float wb8s_at(const Grid &g, const Opts &o, const Consts &c, const ForceIn &in,
                                         int i, int j, int k) {
  const long p = g.at(i, j, k);
  const float qvbr = MPU(CRESS_QVBR, in.qvbr, p, k);
  const float qvd = in.q[p] - qvbr;
  const float rst = MPU(CRESS_RST, in.rst, p, k), ptbr = MPU(CRESS_PTBR, in.ptbr, p, k);
  if (o.cphopt == 0)
    return c.g05 * rst * (qvd / (0.622f + qvbr) - qvd / (1.0f + qvbr) + in.ptp[p] / ptbr);
  // qall: totalqwi.f90:163-164, qc+qr+qi+qs+qg left to right
  const float qall = in.q[p + in.qn] + in.q[p + 2 * in.qn] + in.q[p + 3 * in.qn] + in.q[p + 4 * in.qn] +
                     in.q[p + 5 * in.qn];
  return c.g05 * rst * (qvd / (0.622f + qvbr) - (qvd + qall) / (1.0f + qvbr) + in.ptp[p] / ptbr);
}

// wfrc over i=2..ni-2, j=2..nj-2, k=2..nk-1
template <bool SM, bool U>
__global__ void k_force_w(Tiles T, Grid g, Opts o, Consts c, ForceIn in, float *wfrc) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 2 + xt.y * blockDim.y + threadIdx.y;
  const int k = 2 + xt.z;
  if (i < 2 || i > g.ni - 2 || j > g.nj - 2) return;
  float f = 0.0f;
  if (SM) {
    const bool band = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
    f = smooth4(f, c, o.smtopt, band, k == 2, k == g.nk - 1, [&](int di, int dj, int dk) {
      const long p = g.at(i + di, j + dj, k + dk);
      const int kk = k + dk;
      return MPU(CRESS_RST8W, in.rst8w, p, kk) * in.wp[p] / MPU(CRESS_JCB8W, in.jcb8w, p, kk);
    });
  }
  if (o.vsp_w && k >= o.ksp0 && k <= g.nk - 1)
    f = f - rbct2(in.rbct, g, i, j, k) * MTU(CRESS_RST8W, in.rst8w, i, j, k) * AT(in.wp, i, j, k);
  if (o.buyopt == 1 && k >= 3 && k <= g.nk - 2)
    f = f + (wb8s_at<U>(g, o, c, in, i, j, k - 1) + wb8s_at<U>(g, o, c, in, i, j, k));
  AT(wfrc, i, j, k) = adv_w(f, g, o, c, in.w, in.rstxu, in.rstxv, in.rstxwc, i, j, k);
}

// ---------------------------------------------------------------------------
// Pressure forcing: s_forcep (forcep.f90:254-308) = advp (advp.f90:249-785,
// diaopt=0) then vspp (vspp.f90:180-194).  pfrc over the interior.  No
// numerical smoothing: s_forcedrv passes the constant-zero option id in
// forcep's smtopt slot (forcedrv.f90:598, idzeropt), so the reference's
// large step never smooths p'.
// ---------------------------------------------------------------------------
template <bool U>
__global__ void k_force_p(Tiles T, Grid g, Opts o, Consts c, const float *jcb, const float *jcb8u,
                          const float *jcb8v, const float *jcb8w, const float *u, const float *v,
                          const float *wc, const float *pp, const float *ppp, const float *rbct,
                          float *pfrc) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 2 + xt.y * blockDim.y + threadIdx.y;
  const int k = 2 + xt.z;
  if (i < 2 || i > g.ni - 2 || j > g.nj - 2) return;
  auto jxu = [&](int ii, int jj) { return MTU(CRESS_JCB8U, jcb8u, ii, jj, k) * AT(u, ii, jj, k); };
  auto jxv = [&](int ii, int jj) { return MTU(CRESS_JCB8V, jcb8v, ii, jj, k) * AT(v, ii, jj, k); };
  auto jxw = [&](int kk) { return MTU(CRESS_JCB8W, jcb8w, i, j, kk) * AT(wc, i, j, kk); };
  auto t1 = [&](int ii) { return jxu(ii, j) * (AT(pp, ii, j, k) - AT(pp, ii - 1, j, k)) * c.dxv05n; };
  auto t2 = [&](int jj) { return jxv(i, jj) * (AT(pp, i, jj, k) - AT(pp, i, jj - 1, k)) * c.dyv05n; };
  auto t3 = [&](int kk) { return jxw(kk) * (AT(pp, i, j, kk) - AT(pp, i, j, kk - 1)) * c.dzv05n; };
  float f;
  if (o.advopt == 1) {
    f = (t3(k) + t3(k + 1)) + ((t1(i) + t1(i + 1)) + (t2(j) + t2(j + 1)));
  } else {
    f = (t1(i) + t1(i + 1)) + (t2(j) + t2(j + 1));
    const bool h4 = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
    if (h4) {
      auto T1 = [&](int ii) {
        return (jxu(ii, j) + jxu(ii + 1, j)) * (AT(pp, ii + 1, j, k) - AT(pp, ii - 1, j, k)) * c.dxv24;
      };
      auto T2 = [&](int jj) {
        return (jxv(i, jj) + jxv(i, jj + 1)) * (AT(pp, i, jj + 1, k) - AT(pp, i, jj - 1, k)) * c.dyv24;
      };
      f = c.fourd3 * f + ((T1(i - 1) + T1(i + 1)) + (T2(j - 1) + T2(j + 1)));
    }
    if (o.advopt == 2) {
      f = f + (t3(k) + t3(k + 1));
    } else {
      float vadv = t3(k) + t3(k + 1);
      if (h4 && k >= 3 && k <= g.nk - 3) {
        auto T3 = [&](int kk) {
          return (jxw(kk) + jxw(kk + 1)) * (AT(pp, i, j, kk + 1) - AT(pp, i, j, kk - 1)) * c.dzv24;
        };
        vadv = c.fourd3 * vadv + (T3(k - 1) + T3(k + 1));
      }
      f = f + vadv;
    }
  }
  if (o.vsp_p && k >= o.ksp0 - 1 && k <= g.nk - 2)
    f = f - 0.5f * (rbct2(rbct, g, i, j, k) + rbct2(rbct, g, i, j, k + 1)) * MTU(CRESS_JCB, jcb, i, j, k) *
                AT(ppp, i, j, k);
  AT(pfrc, i, j, k) = f;
}

// ---------------------------------------------------------------------------
// Scalar forcing: s_forcept/s_forceqv/s_forceq (forcept.f90:344-438,
// forceqv.f90:290-372, forceq.f90:394-...) = advs (advs.f90:205-445) then the
// vspopt=2 sponge (vsps.f90:183-197, vspqv.f90:185-200, vsps0.f90:114-127)
// and, for theta', advbspt (advbspt.f90:143-190).  Scalars: 0 ptp, 1..6
// qv..qg, 7 nci.
// ---------------------------------------------------------------------------


__device__ float adv_s(const Grid &g, const Opts &o, const Consts &c, const float *s, const float *rxu,
                       const float *rxv, const float *rxw, int i, int j, int k) {
  auto t1 = [&](int ii) { return AT(rxu, ii, j, k) * (AT(s, ii, j, k) - AT(s, ii - 1, j, k)) * c.dxv05n; };
  auto t2 = [&](int jj) { return AT(rxv, i, jj, k) * (AT(s, i, jj, k) - AT(s, i, jj - 1, k)) * c.dyv05n; };
  auto t3 = [&](int kk) { return AT(rxw, i, j, kk) * (AT(s, i, j, kk) - AT(s, i, j, kk - 1)) * c.dzv05n; };
  if (o.advopt == 1) return (t3(k) + t3(k + 1)) + ((t1(i) + t1(i + 1)) + (t2(j) + t2(j + 1)));
  float f = (t1(i) + t1(i + 1)) + (t2(j) + t2(j + 1));
  const bool h4 = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
  if (h4) {
    auto T1 = [&](int ii) {
      return (AT(rxu, ii, j, k) + AT(rxu, ii + 1, j, k)) * (AT(s, ii + 1, j, k) - AT(s, ii - 1, j, k)) * c.dxv24;
    };
    auto T2 = [&](int jj) {
      return (AT(rxv, i, jj, k) + AT(rxv, i, jj + 1, k)) * (AT(s, i, jj + 1, k) - AT(s, i, jj - 1, k)) * c.dyv24;
    };
    f = c.fourd3 * f + ((T1(i - 1) + T1(i + 1)) + (T2(j - 1) + T2(j + 1)));
  }
  if (o.advopt == 2) return f + (t3(k) + t3(k + 1));
  float vadv = t3(k) + t3(k + 1);
  if (h4 && k >= 3 && k <= g.nk - 3) {
    auto T3 = [&](int kk) {
      return (AT(rxw, i, j, kk) + AT(rxw, i, j, kk + 1)) * (AT(s, i, j, kk + 1) - AT(s, i, j, kk - 1)) * c.dzv24;
    };
    vadv = c.fourd3 * vadv + (T3(k - 1) + T3(k + 1));
  }
  return f + vadv;
}

// The stencil values of one scalar that advs reads at (i,j,k): the cell,
// its neighbours at distance 1 in x, y, z, and at distance 2 where the
// 4th-order terms apply (h4: x and y, v4: z).
struct SStencil {
  float s0, xm1, xp1, ym1, yp1, zm1, zp1, xm2, xp2, ym2, yp2, zm2, zp2;
};

__device__ __forceinline__ void adv_s_load(const Grid &g, const Opts &o, bool h4, bool v4, const float *s,
                                           uint32_t q, SStencil &L) {
  const uint32_t sx = 4u * (uint32_t)g.nx, sk = 4u * (uint32_t)g.nxy;
  L.s0 = LD(s, q);
  L.xm1 = LD(s, q - 4);
  L.xp1 = LD(s, q + 4);
  L.ym1 = LD(s, q - sx);
  L.yp1 = LD(s, q + sx);
  L.zm1 = LD(s, q - sk);
  L.zp1 = LD(s, q + sk);
  if (o.advopt != 1 && h4) {
    L.xm2 = LD(s, q - 8);
    L.xp2 = LD(s, q + 8);
    L.ym2 = LD(s, q - 2 * sx);
    L.yp2 = LD(s, q + 2 * sx);
    if (o.advopt == 3 && v4) {
      L.zm2 = LD(s, q - 2 * sk);
      L.zp2 = LD(s, q + 2 * sk);
    }
  }
}

// advs for one scalar at (i,j,k) from flux weights and stencil values
// already in registers: fu[m] = rstxu(i-1+m), fv[m] = rstxv(j-1+m),
// fw[m] = rstxwc(k-1+m), m=0..3.  Same terms and operation order as adv_s.
__device__ __forceinline__ float adv_s_val(const Opts &o, const Consts &c, bool h4, bool v4, const float *fu,
                                           const float *fv, const float *fw, const SStencil &L) {
  const float s0 = L.s0;
  const float sxm1 = L.xm1, sxp1 = L.xp1;
  const float sym1 = L.ym1, syp1 = L.yp1;
  const float szm1 = L.zm1, szp1 = L.zp1;
  const float t1a = fu[1] * (s0 - sxm1) * c.dxv05n, t1b = fu[2] * (sxp1 - s0) * c.dxv05n;
  const float t2a = fv[1] * (s0 - sym1) * c.dyv05n, t2b = fv[2] * (syp1 - s0) * c.dyv05n;
  const float t3a = fw[1] * (s0 - szm1) * c.dzv05n, t3b = fw[2] * (szp1 - s0) * c.dzv05n;
  if (o.advopt == 1) return (t3a + t3b) + ((t1a + t1b) + (t2a + t2b));
  float f = (t1a + t1b) + (t2a + t2b);
  if (h4) {
    const float sxm2 = L.xm2, sxp2 = L.xp2;
    const float sym2 = L.ym2, syp2 = L.yp2;
    const float T1a = (fu[0] + fu[1]) * (s0 - sxm2) * c.dxv24;
    const float T1b = (fu[2] + fu[3]) * (sxp2 - s0) * c.dxv24;
    const float T2a = (fv[0] + fv[1]) * (s0 - sym2) * c.dyv24;
    const float T2b = (fv[2] + fv[3]) * (syp2 - s0) * c.dyv24;
    f = c.fourd3 * f + ((T1a + T1b) + (T2a + T2b));
  }
  if (o.advopt == 2) return f + (t3a + t3b);
  float vadv = t3a + t3b;
  if (h4 && v4) {
    const float szm2 = L.zm2, szp2 = L.zp2;
    const float T3a = (fw[0] + fw[1]) * (s0 - szm2) * c.dzv24;
    const float T3b = (fw[2] + fw[3]) * (szp2 - s0) * c.dzv24;
    vadv = c.fourd3 * vadv + (T3a + T3b);
  }
  return f + vadv;
}

// forcept/forceqv/forceq for all scalars of one interior cell: advs
// (advs.f90:205-420), the vspopt=2 sponge (vsps.f90:160-198, vspqv) and, for
// theta', advbspt (advbspt.f90:143-190).  The flux weights are loaded once
// and shared by the scalars.
template <bool SM, bool U>
__global__ void __launch_bounds__(64 * CRESS_FS_BYMAX) k_force_s(Tiles T, Grid g, Opts o, Consts c, ScalarIn in, int nscalar) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 2 + xt.y * blockDim.y + threadIdx.y;
  const int k = 2 + xt.z;
  if (i < 2 || i > g.ni - 2 || j > g.nj - 2) return;
  const uint32_t sx = 4u * (uint32_t)g.nx, sk = 4u * (uint32_t)g.nxy;
  const uint32_t q = 4u * (uint32_t)g.at(i, j, k);
  float fu[4], fv[4], fw[4];
  for (int m = 0; m < 4; m++) {
    fu[m] = LD(in.rstxu, q + 4u * m - 4u);
    fv[m] = LD(in.rstxv, q + sx * m - sx);
  }
  fw[0] = LD(in.rstxwc, q - sk);
  fw[1] = LD(in.rstxwc, q);
  fw[2] = LD(in.rstxwc, q + sk);
  fw[3] = LD(in.rstxwc, q + 2 * sk);
  const bool h4 = j >= 2 + o.js && j <= g.nj - 2 - o.jn && i >= 2 + o.iw && i <= g.ni - 2 - o.ie;
  const bool v4 = k >= 3 && k <= g.nk - 3;
  const bool sponge = k >= o.ksp0 - 1 && k <= g.nk - 2;
  float r = 0.0f, rst = 0.0f;
  if (sponge) {
    r = 0.5f * (rbct2(in.rbct, g, i, j, k) + rbct2(in.rbct, g, i, j, k + 1));
    rst = MQU(CRESS_RST, in.rst, q, k);
  }
  // each scalar's stencil is loaded before the previous scalar's forcing
  // is stored, so that the loads of consecutive scalars overlap
  SStencil cur, nxt;
  adv_s_load(g, o, h4, v4, in.s[0], q, cur);
  for (int sc = 0; sc < nscalar; sc++) {
    if (sc + 1 < nscalar) adv_s_load(g, o, h4, v4, in.s[sc + 1], q, nxt);
    float f = adv_s_val(o, c, h4, v4, fu, fv, fw, cur);
    cur = nxt;
    if (SM) {
      const float *sp = in.sp[sc];
      f = smooth4(f, c, o.smtopt, h4, k == 2, k == g.nk - 2, [&](int di, int dj, int dk) {
        const long p = g.at(i + di, j + dj, k + dk);
        const int kk = k + dk;
        return sc == 1 ? MPU(CRESS_RBR, in.rbr, p, kk) * (sp[p] - MPU(CRESS_QVBR, in.qvbr, p, kk))
                       : MPU(CRESS_RBR, in.rbr, p, kk) * sp[p];
      });
    }
    const int vsp = sc == 0 ? o.vsp_pt : (sc == 1 ? o.vsp_qv : o.vsp_q);
    if (vsp && sponge) {
      const float d = sc == 1 ? LD(in.sp[sc], q) - MQU(CRESS_QVBR, in.qvbr, q, k) : LD(in.sp[sc], q);
      f = f - r * rst * d;
    }
    if (sc == 0) {
      const float dziv25 = 0.25f * c.dziv;
      // level kk of offset qq
      auto a8 = [&](uint32_t qq, int kk) {
        return (MQU(CRESS_RBR, in.rbr, qq - sk, kk - 1) + MQU(CRESS_RBR, in.rbr, qq, kk)) * LD(in.w, qq) *
               (MQU(CRESS_PTBR, in.ptbr, qq - sk, kk - 1) - MQU(CRESS_PTBR, in.ptbr, qq, kk)) * dziv25;
      };
      f = f + (a8(q, k) + a8(q + sk, k + 1));
    }
    ST(in.frc[sc], q, f);
  }
}

// ---------------------------------------------------------------------------
// Acoustic substep (s_hevi, hevi.f90:339-427).
// ---------------------------------------------------------------------------

// pgrad.f90:260-306 with diver3d.f90:182-326 (the divergence-damped pressure
// ptil = pp + divcv*jcb*div3d over i=1..ni-1, j=1..nj-1, k=1..nk-1) fused
// with stepuv.f90:288-310's interior update uf += dts*(ufrc+upg)/rst8u,
// upg = ((jcb*ptil)(i-1) - (jcb*ptil)(i))/dx (pgrad.f90:357-395), and the
// same for vf.  A 64 x blockDim.y tile computes ptil on itself and on its i-1 column
// and j-1 row into LDS, then updates uf and vf from it.  uf and vf are read
// from one buffer and written to another, because neighbouring tiles read
// the old values for their halo ptil; ptil goes to memory for the column
// kernel's vertical pressure gradient.
template <bool U>
__device__ __forceinline__ float ptil_at(const Grid &g, const Opts &o, const Consts &c, const AcPtUv &a, int i,
                                         int j, int k) {
  const uint32_t q = g.off(i, j, k), sx = 4u * (uint32_t)g.nx, sk = 4u * (uint32_t)g.nxy;
  if (o.divopt == 0) return LD(a.ppf, q);
  const float s1a = MQU(CRESS_RST8U, a.rst8u, q, k) * LD(a.uf, q);
  const float s1b = MQU(CRESS_RST8U, a.rst8u, q + 4, k) * LD(a.uf, q + 4);
  const float s2a = MQU(CRESS_RST8V, a.rst8v, q, k) * LD(a.vf, q);
  const float s2b = MQU(CRESS_RST8V, a.rst8v, q + sx, k) * LD(a.vf, q + sx);
  const float s3a = MQU(CRESS_RST8W, a.rst8w, q, k) * LD(a.wc, q);
  const float s3b = MQU(CRESS_RST8W, a.rst8w, q + sk, k + 1) * LD(a.wc, q + sk);
  const float div = (s3a - s3b) * c.dziv + ((s1a - s1b) * c.dxiv + (s2a - s2b) * c.dyiv);
  return LD(a.ppf, q) + a.divcv * MQU(CRESS_JCB, a.jcb, q, k) * div;
}

template <bool U, int SLOTS>
__global__ void __launch_bounds__(64 * CRESS_PTUV_BYMAX) k_ac_ptuv(Tiles T, Grid g, Opts o, Consts c, AcPtUv a) {
  const uint3 xt = xcd_tile(T);
  __shared__ float P[CRESS_PTUV_BYMAX + 1][66];   // P[y][x]: j = j0-1+y, i = i0-1+x
  const int tx = threadIdx.x, ty = threadIdx.y, by = blockDim.y;
  const int i0 = xt.x * 64, j0 = 1 + xt.y * by, k = 1 + xt.z;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i >= 1 && i <= g.ni - 1 && j <= g.nj - 1;
  const uint32_t q = g.off(i, j, k), sx = 4u * (uint32_t)g.nx;
  // the update's own-cell operands are loaded first: they do not depend on
  // the shared ptil values
  const bool upd = own && k >= 2 && k <= g.nk - 2 && i >= 2 && j >= 2;
  const bool du = upd && j <= g.nj - 2, dv = upd && i <= g.ni - 2;
  float uf0 = 0.0f, ufr = 0.0f, vf0 = 0.0f, vfr = 0.0f;
  if (du) {
    uf0 = LD(a.uf, q);
    ufr = LD(a.ufrc, q);
  }
  if (dv) {
    vf0 = LD(a.vf, q);
    vfr = LD(a.vfrc, q);
  }
  float pown;
  if (SLOTS) {
    // ptil on the (by+1) x 65 points of the tile with its i-1 column and j-1
    // row, in slots of consecutive points per thread, all loads issued
    // without branching on the thread's position
    const int np = (by + 1) * 65, tid = ty * 64 + tx;
    for (int e = tid; e < np; e += 64 * by) {
      const int yy = e / 65, xx = e - 65 * (e / 65);
      const int ii = i0 - 1 + xx, jj = j0 - 1 + yy;
      const bool ok = ii >= 1 && ii <= g.ni - 1 && jj >= 1 && jj <= g.nj - 1;
      P[yy][xx] = ok ? ptil_at<U>(g, o, c, a, ii, jj, k) : 0.0f;
    }
    __syncthreads();
    if (!own) return;
    pown = P[ty + 1][tx + 1];
  } else {
    // each thread its own point; the first row and the first lane also the
    // halo row and column
    pown = own ? ptil_at<U>(g, o, c, a, i, j, k) : 0.0f;
    P[ty + 1][tx + 1] = pown;
    if (ty == 0) P[0][tx + 1] = (j0 - 1 >= 1 && i >= 1 && i <= g.ni - 1) ? ptil_at<U>(g, o, c, a, i, j0 - 1, k) : 0.0f;
    if (tx == 0) P[ty + 1][0] = (i0 - 1 >= 1 && j <= g.nj - 1) ? ptil_at<U>(g, o, c, a, i0 - 1, j, k) : 0.0f;
    __syncthreads();
    if (!own) return;
  }
  ST(a.ptil, q, pown);
  if (!upd) return;
  const float pc = MQU(CRESS_JCB, a.jcb, q, k) * pown;
  if (du) {
    const float upg = (MQU(CRESS_JCB, a.jcb, q - 4, k) * P[ty + 1][tx] - pc) * c.dxiv;
    ST(a.ufn, q, uf0 + a.dts * (ufr + upg) / MQU(CRESS_RST8U, a.rst8u, q, k));
  }
  if (dv) {
    const float vpg = (MQU(CRESS_JCB, a.jcb, q - sx, k) * P[ty][tx + 1] - pc) * c.dyiv;
    ST(a.vfn, q, vf0 + a.dts * (vfr + vpg) / MQU(CRESS_RST8V, a.rst8v, q, k));
  }
}

// The vertical part of the substep for one interior column (i,j):
// diverpih (diverpih.f90:188-214), diverpiv, advbspi 'back', pgradiv
// (pgradiv.f90:143-170), buoywsi (buoywsi.f90:165-211), stepwi
// (stepwi.f90:319-427) with the Thomas solve of gaussel (gaussel.f90:169-262),
// phy2cnt+vbcwc+vbcw for this column, diverpiv + advbspi 'fore', steppi
// (steppi.f90:211-223) and vbcp (vbcp.f90:126-162).
//
// One lane per column.  Two per-column intermediates go to global slots
// that nothing else reads while the kernel runs: fp(k) into wc(k) (wc is
// written with its final value in the last pass), and the Thomas
// super-diagonal ss'(k) into ptil(k-1), whose value the column has consumed
// by then (ptil is read by this column only, and recomputed before its
// next use).  ss' depends on the metrics and dts alone, so when those are
// horizontally uniform it is one number per level: the workgroup's first
// lane keeps it in an LDS table instead, and the ptil round trip goes away.
template <bool UNI>
__global__ void __launch_bounds__(64 * CRESS_COL_BYMAX) k_ac_col(Tiles T, Grid g, Opts o, Consts c, AcCol a) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 2 + xt.y * blockDim.y + threadIdx.y;
  const bool act = i >= 2 && i <= g.ni - 2 && j <= g.nj - 2;
  // UNI: the metrics are horizontally uniform, and the Thomas coefficients
  // are read from a per-level table the host formed with this kernel's
  // operations (AcCol::coef)
  constexpr bool uni = UNI, U = UNI;
  const float *coef = a.coef;
  const int nk = g.nk;
  // byte offsets (a field is below 4 GiB; checked at context creation)
  const uint32_t sk = 4u * (uint32_t)g.nxy, sx = 4u * (uint32_t)g.nx;
  float *ptil = a.ptil;
  float *wf = a.wf, *wc = a.wc, *ppf = a.ppf;
  const float dts = a.dts;
  const float dtdzw = dts * c.dziv * c.weicoe;
  const float dtw = dts * c.weicoe;
  const float sbsqzi = dts * dts * c.weicoe * c.weicoe * c.dziv;
  const float sbsqg5 = c.g05 * dts * dts * c.weicoe * c.weicoe;

  // pass 1 (upward, k=2..nk-2): fp, fw, explicit w, tridiagonal forward sweep
  float fpdvj_km1 = 0.0f, wb_km1 = 0.0f, pt_km1 = 0.0f;
  float fwc_km1 = 0.0f, wcc_km1 = 0.0f, rcsq_km1 = 0.0f, rst_km1 = 0.0f;
  float ssn_km1 = 0.0f, ff_km1 = 0.0f;
  const uint32_t q0 = 4u * (uint32_t)g.at(i, j, 2);
  float wf_k = act ? LD(wf, q0) : 0.0f;
  uint32_t q = q0;
  // The field values of level k+1 are loaded before level k's stores, so
  // that their latency overlaps level k's arithmetic (the loads of one
  // level do not wait for the previous level's writes).  Level k+1 reads
  // wf(k+2), ptil(k+1) and ppf(k+1), which pass 1 has not written yet.
  float n_wf1 = 0.0f, n_u0 = 0.0f, n_u1 = 0.0f, n_v0 = 0.0f, n_v1 = 0.0f;
  float n_pf = 0.0f, n_pt = 0.0f, n_pp = 0.0f, n_fw = 0.0f;
  auto pre1 = [&](uint32_t qq) {
    n_wf1 = LD(wf, qq + sk);
    n_u0 = LD(a.uf, qq);
    n_u1 = LD(a.uf, qq + 4);
    n_v0 = LD(a.vf, qq);
    n_v1 = LD(a.vf, qq + sx);
    n_pf = LD(a.pfrc, qq);
    n_pt = LD(ptil, qq);
    n_pp = LD(ppf, qq);
    n_fw = LD(a.wfrc, qq);
  };
  if (act) pre1(q0);
  for (int k = 2; k <= nk - 2 && act; k++, q += sk) {
    const float wf_kp1 = n_wf1, u0 = n_u0, u1 = n_u1, v0 = n_v0, v1 = n_v1;
    const float pfrc = n_pf, pt_k = n_pt, pp_k = n_pp, wfrc = n_fw;
    if (k + 1 <= nk - 2) pre1(q + sk);
    const float rcsq = MQU(CRESS_RCSQ, a.rcsq, q, k), rst = MQU(CRESS_RST, a.rst, q, k);
    const float jcb = MQU(CRESS_JCB, a.jcb, q, k), rbr = MQU(CRESS_RBR, a.rbr, q, k);
    const float ju0 = MQU(CRESS_JCB8U, a.jcb8u, q, k) * u0;
    const float ju1 = MQU(CRESS_JCB8U, a.jcb8u, q + 4, k) * u1;
    const float jv0 = MQU(CRESS_JCB8V, a.jcb8v, q, k) * v0;
    const float jv1 = MQU(CRESS_JCB8V, a.jcb8v, q + sx, k) * v1;
    const float hdiv = rcsq * ((ju0 - ju1) * c.dxiv + (jv0 - jv1) * c.dyiv);
    const float vdiv = rcsq * (wf_k - wf_kp1) * c.dziv;
    const float fp = pfrc + hdiv + c.weic1m * (c.g05 * (wf_k + wf_kp1) * rst + vdiv);
    ST(wc, q, fp);
    const float fpdvj = fp / jcb;
    const float wb = (pp_k * rst + fp * rbr * dtw) / rcsq * c.g05n;
    float fwc = 0.0f, wcc = 0.0f;
    if constexpr (!UNI) {
      const float t1 = c.g05 * rbr / rcsq;
      const float t2 = c.dziv / jcb;
      fwc = t1 + t2;
      wcc = t1 - t2;
    }
    if (k >= 3 && uni) {
      const float wpg = (pt_km1 - pt_k) * c.dziv;
      float fw = wfrc + wpg + (fpdvj_km1 - fpdvj) * dtdzw;
      if (o.buyopt == 1) fw = fw + (wb_km1 + wb);
      const float rst8w = MTAB(CRESS_RST8W, k);
      float wfn = wf_k + dts * fw / rst8w;
      const float rr = coef[3 * k], r = coef[3 * k + 1];
      if (k == 3) wfn = wfn + 0.25f * rr * (0.0f + 0.0f);   // bbcw on flat terrain
      const float ffn = k == 3 ? wfn * r : (wfn - rr * ff_km1) * r;
      ST(wf, q, ffn);
      ff_km1 = ffn;
    } else if (k >= 3) {
      const float wpg = (pt_km1 - pt_k) * c.dziv;
      float fw = wfrc + wpg + (fpdvj_km1 - fpdvj) * dtdzw;
      if (o.buyopt == 1) fw = fw + (wb_km1 + wb);
      const float rst8w = MQU(CRESS_RST8W, a.rst8w, q, k);
      float wfn = wf_k + dts * fw / rst8w;
      float rr, ss, tt;
      if (o.buyopt == 0) {
        const float aa = sbsqzi / rst8w;
        const float mm = aa * rcsq_km1;
        const float nn = aa * rcsq;
        rr = -mm * fwc_km1;
        ss = 1.0f + (nn * fwc - mm * wcc_km1);
        tt = nn * wcc;
      } else {
        const float rstiv = 1.0f / rst8w;
        const float aa = rstiv * sbsqzi;
        const float bb = rstiv * sbsqg5;
        const float mm = bb * rst_km1 - aa * rcsq_km1;
        const float nn = bb * rst + aa * rcsq;
        rr = mm * fwc_km1;
        ss = 1.0f + (mm * wcc_km1 + nn * fwc);
        tt = nn * wcc;
      }
      if (k == 3) wfn = wfn + 0.25f * rr * (0.0f + 0.0f);   // bbcw on flat terrain
      // Thomas forward sweep (gaussel.f90:189-236), kstr=3, kend=nk-2
      float ssn, ffn;
      if (k == 3) {
        const float r = 1.0f / ss;
        ssn = tt * r;
        ffn = wfn * r;
      } else {
        const float r = 1.0f / (ss - rr * ssn_km1);
        ssn = tt * r;
        ffn = (wfn - rr * ff_km1) * r;
      }
      ST(ptil, q - sk, ssn);
      ST(wf, q, ffn);
      ssn_km1 = ssn;
      ff_km1 = ffn;
    }
    fpdvj_km1 = fpdvj;
    wb_km1 = wb;
    fwc_km1 = fwc;
    wcc_km1 = wcc;
    rcsq_km1 = rcsq;
    rst_km1 = rst;
    pt_km1 = pt_k;
    wf_k = wf_kp1;
  }
  if (!act) return;
  // pass 2 (downward, k=nk-2..2): back substitution (gaussel.f90:248-260)
  // fused with diverpiv + advbspi 'fore' and steppi, whose per-level update
  // of ppf needs only w(k), w(k+1) and fp(k), in any order of k.  The
  // levels 2 and nk-1 take w after vbcw; wc gets its final value.
  const float wc2 = wc_value<U>(g, o, wf, a.jcb8w, i, j, 2);
  const float wcm = wc_value<U>(g, o, wf, a.jcb8w, i, j, nk - 1);
  const float w2 = o.bbc == 2 ? -0.25f * ((0.0f + 0.0f) + (0.0f + 0.0f)) : AT(wf, i, j, 2);
  const float wm = o.tbc == 2 ? 0.0f : AT(wf, i, j, nk - 1);
  float wkp1 = wm;
  float wk = ff_km1;   // wf(nk-2): the last forward-sweep value is final
  q = q0 + (uint32_t)(nk - 4) * sk;
  // as in pass 1, level k-1's values are loaded before level k's stores
  // (pass 2 writes each level once, going down)
  float m_wf = 0.0f, m_ss = 0.0f, m_fp = 0.0f, m_pp = 0.0f;
  auto pre2 = [&](int kk, uint32_t qq) {
    if (kk <= nk - 3 && kk >= 3) {
      m_wf = LD(wf, qq);
      m_ss = uni ? coef[3 * kk + 2] : LD(ptil, qq - sk);
    }
    m_fp = LD(wc, qq);
    m_pp = LD(ppf, qq);
  };
  pre2(nk - 2, q);
  for (int k = nk - 2; k >= 2; k--, q -= sk) {
    const float ff_k = m_wf, ss_k = m_ss, fp_k = m_fp, pp_k = m_pp;
    if (k - 1 >= 2) pre2(k - 1, q - sk);
    if (k <= nk - 3 && k >= 3) {
      wk = ff_k - ss_k * wkp1;
      ST(wf, q, wk);
    } else if (k == 2) {
      wk = w2;
    }
    const float vdiv = MQU(CRESS_RCSQ, a.rcsq, q, k) * (wk - wkp1) * c.dziv;
    const float fp = fp_k + c.weicoe * (c.g05 * (wk + wkp1) * MQU(CRESS_RST, a.rst, q, k) + vdiv);
    ST(ppf, q, pp_k + dts * fp / MQU(CRESS_JCB, a.jcb, q, k));
    ST(wc, q, k == 2 ? wc2 : (o.sthopt == 0 ? wk : wk / MQU(CRESS_JCB8W, a.jcb8w, q, k)));
    wkp1 = wk;
  }
  // phy2cnt+vbcwc at the outer rows and vbcw, from the solved column
  AT(wc, i, j, nk - 1) = wcm;
  AT(wc, i, j, 1) = wc_value<U>(g, o, wf, a.jcb8w, i, j, 1);
  AT(wc, i, j, nk) = wc_value<U>(g, o, wf, a.jcb8w, i, j, nk);
  if (o.bbc == 2) {
    const float z = 0.25f * ((0.0f + 0.0f) + (0.0f + 0.0f));
    AT(wf, i, j, 1) = -MT(CRESS_JCB8W, a.jcb8w, i, j, 3) * wc_raw<U>(g, o.sthopt, wf, a.jcb8w, i, j, 3) - z;
    AT(wf, i, j, 2) = w2;
  } else {
    AT(wf, i, j, 1) = AT(wf, i, j, 2);
  }
  if (o.tbc == 2) {
    const float z = 0.25f * ((0.0f + 0.0f) + (0.0f + 0.0f));
    AT(wf, i, j, nk) = -MT(CRESS_JCB8W, a.jcb8w, i, j, nk - 2) * wc_raw<U>(g, o.sthopt, wf, a.jcb8w, i, j, nk - 2) - z;
    AT(wf, i, j, nk - 1) = 0.0f;
  } else {
    AT(wf, i, j, nk) = AT(wf, i, j, nk - 1);
  }
  // vbcp for this column
  if (o.bbc == 2) AT(ppf, i, j, 1) = 2.0f * AT(ppf, i, j, 2) - AT(ppf, i, j, 3);
  else if (o.bbc == 3) AT(ppf, i, j, 1) = AT(ppf, i, j, 2);
  AT(ppf, i, j, nk - 1) = AT(ppf, i, j, nk - 2);
}

// Ring columns (i or j on the first halo row, within 1..ni-1 x 1..nj-1)
// after the lateral fill of wf and ppf: phy2cnt+vbcwc, vbcw and vbcp exactly
// as the reference applies them over i=1..ni-1, j=1..nj-1 (stepwi.f90:540-549,
// steppi.f90:316).  One thread per (ring column, level): every value a level
// needs is recomputed from wf/ppf levels that no thread of the launch writes
// (wc at k from w(k); the bottom and top rows from levels 2, 3, nk-2, nk-1).
__global__ void k_ac_ring(Grid g, Opts o, const float *jcb8w, float *wf, float *wc, float *ppf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.y;
  const int nxr = g.ni - 1, nyr = g.nj - 1;   // ring of the (1..ni-1)x(1..nj-1) box
  const int nring = 2 * nxr + 2 * (nyr - 2);
  if (t >= nring || k > g.nk) return;
  int i, j;
  if (t < nxr) { i = 1 + t; j = 1; }
  else if (t < 2 * nxr) { i = 1 + (t - nxr); j = g.nj - 1; }
  else if (t < 2 * nxr + (nyr - 2)) { i = 1; j = 2 + (t - 2 * nxr); }
  else { i = g.ni - 1; j = 2 + (t - 2 * nxr - (nyr - 2)); }
  const int nk = g.nk;
  auto raw = [&](int kk) { return wc_raw<false>(g, o.sthopt, wf, jcb8w, i, j, kk); };
  const float z = 0.25f * ((0.0f + 0.0f) + (0.0f + 0.0f));
  if (k == 1) {
    if (o.bbc == 2) {
      const float w3 = raw(3);
      AT(wc, i, j, 1) = -w3;
      AT(wf, i, j, 1) = -MT(CRESS_JCB8W, jcb8w, i, j, 3) * w3 - z;
      AT(ppf, i, j, 1) = 2.0f * AT(ppf, i, j, 2) - AT(ppf, i, j, 3);
    } else {
      AT(wc, i, j, 1) = raw(2);
      AT(wf, i, j, 1) = AT(wf, i, j, 2);
      if (o.bbc == 3) AT(ppf, i, j, 1) = AT(ppf, i, j, 2);
    }
  } else if (k == nk) {
    if (o.tbc == 2) {
      const float wm = raw(nk - 2);
      AT(wc, i, j, nk) = -wm;
      AT(wf, i, j, nk) = -MT(CRESS_JCB8W, jcb8w, i, j, nk - 2) * wm - z;
    } else {
      AT(wc, i, j, nk) = raw(nk - 1);
      AT(wf, i, j, nk) = AT(wf, i, j, nk - 1);
    }
  } else if (k == 2 && o.bbc == 2) {
    AT(wc, i, j, 2) = 0.0f;
    AT(wf, i, j, 2) = -0.25f * ((0.0f + 0.0f) + (0.0f + 0.0f));
  } else if (k == nk - 1 && o.tbc == 2) {
    AT(wc, i, j, nk - 1) = 0.0f;
    AT(wf, i, j, nk - 1) = 0.0f;
    AT(ppf, i, j, nk - 1) = AT(ppf, i, j, nk - 2);
  } else {
    AT(wc, i, j, k) = raw(k);
    if (k == nk - 1) AT(ppf, i, j, nk - 1) = AT(ppf, i, j, nk - 2);
  }
}

// ---------------------------------------------------------------------------
// Scalar leapfrog, s_steps (steps.f90:520-760): s_f = max(0, s_p + frc*dtdrst)
// for qv..qg, without the clamp for theta' and nci; dtdrst = 2*dtb/rst.
// The q and nci forcings are read from the future arrays themselves.
// ---------------------------------------------------------------------------
__global__ void k_steps(Tiles T, Grid g, float dtb2, const float *rst, StepsIn s, int withq) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 2 + xt.y * blockDim.y + threadIdx.y;
  const int k = 2 + xt.z;
  if (i < 2 || i > g.ni - 2 || j > g.nj - 2) return;
  const long p = g.at(i, j, k);
  const float d = dtb2 / MP(CRESS_RST, rst, p, k);
  s.ptpf[p] = s.ptpp[p] + s.ptfrc[p] * d;
  const float a = s.qp[p] + s.qvfrc[p] * d;
  s.qf[p] = a > 0.0f ? a : 0.0f;
  if (!withq) return;
  for (int sp = 1; sp < 6; sp++) {
    const long q = p + sp * s.qn;
    const float b = s.qp[q] + s.qf[q] * d;
    s.qf[q] = b > 0.0f ? b : 0.0f;
  }
  s.ncif[p] = s.ncip[p] + s.ncif[p] * d;
}

// steps.f90's dtdrst = 2*dtb/rst over the interior (the array s_steps
// returns beside the future level).
__global__ void k_dtdrst(Grid g, float dtb2, const float *rst, float *dtdrst) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > g.ni - 2 || j > g.nj - 2) return;
  const long p = g.at(i, j, k);
  dtdrst[p] = dtb2 / MP(CRESS_RST, rst, p, k);
}

// The TKE line of s_steps for tubopt>=2 (steps.f90:997-1012): tkef holds
// the forcing on entry and the future TKE, clipped at zero, on return;
// 2 dtb / rst is formed as k_dtdrst forms it.
__global__ void k_steps_tke(Grid g, float dtb2, const float *rst, const float *tkep, float *tkef) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > g.ni - 2 || j > g.nj - 2) return;
  const long p = g.at(i, j, k);
  const float d = dtb2 / MP(CRESS_RST, rst, p, k);
  const float a = tkep[p] + tkef[p] * d;
  tkef[p] = a > 0.0f ? a : 0.0f;
}

// ---------------------------------------------------------------------------
// Halo fills.
// s_bcycle (bcycle.f90:159-225): x pass over j=0..nj+1 then y pass over
// i=0..ni+1, for up to 16 fields per launch.
// ---------------------------------------------------------------------------
__global__ void k_bcycle_x(Grid g, FillSet fs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.y;
  const int f = blockIdx.z;
  if (j > g.nj + 1 || f >= fs.n) return;
  const FillSpec &s = fs.f[f];
  if (k > s.kmax) return;
  float *v = s.ptr;
  AT(v, s.iwrcv, j, k) = AT(v, s.iesnd, j, k);
  AT(v, s.iercv, j, k) = AT(v, s.iwsnd, j, k);
}

__global__ void k_bcycle_y(Grid g, FillSet fs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.y;
  const int f = blockIdx.z;
  if (i > g.ni + 1 || f >= fs.n) return;
  const FillSpec &s = fs.f[f];
  if (k > s.kmax) return;
  float *v = s.ptr;
  AT(v, i, s.jsrcv, k) = AT(v, i, s.jnsnd, k);
  AT(v, i, s.jnrcv, k) = AT(v, i, s.jssnd, k);
}

// Vertical fills: vbcs (vbcs.f90:116-134), vbcu, vbcv (same with i to ni or
// j to nj), vbcp (vbcp.f90:126-162); kind 0 s, 1 u, 2 v, 3 p.
__global__ void k_vbc(Grid g, int bbc, VbcSet vs) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int f = blockIdx.z;
  if (f >= vs.n) return;
  const int kind = vs.kind[f];
  const int ihi = kind == 1 ? g.ni : g.ni - 1;
  const int jhi = kind == 2 ? g.nj : g.nj - 1;
  if (i > ihi || j > jhi) return;
  float *v = vs.ptr[f];
  const int nk = g.nk;
  if (kind == 3) {
    if (bbc == 2) AT(v, i, j, 1) = 2.0f * AT(v, i, j, 2) - AT(v, i, j, 3);
    else if (bbc == 3) AT(v, i, j, 1) = AT(v, i, j, 2);
  } else {
    AT(v, i, j, 1) = AT(v, i, j, 2);
  }
  AT(v, i, j, nk - 1) = AT(v, i, j, nk - 2);
}

// ---------------------------------------------------------------------------
// Asselin filter, s_timeflt (timeflt.f90:335-560): x = m1fc2*x +
// filcoe*(xf + xp) over each field's own box (u: i=1..ni, j=1..nj-1,
// k=1..nk-1; v: j to nj; w: k to nk; scalars i=1..ni-1, j=1..nj-1).
// ---------------------------------------------------------------------------
__global__ void k_timeflt(Tiles T, Grid g, Consts c, AsselinSet as) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 1 + xt.y * blockDim.y + threadIdx.y;
  const int f = xt.z / g.nk;
  const int k = 1 + xt.z % g.nk;
  if (f >= as.n) return;
  const int kind = as.kind[f];
  const int ihi = kind == 1 ? g.ni : g.ni - 1;
  const int jhi = kind == 2 ? g.nj : g.nj - 1;
  const int khi = kind == 3 ? g.nk : g.nk - 1;
  if (i < 1 || i > ihi || j > jhi || k > khi) return;
  const long p = g.at(i, j, k);
  float *x = as.x[f];
  x[p] = c.m1fc2 * x[p] + c.filcoe * (as.xf[f][p] + as.xp[f][p]);
}

}  // namespace cress

namespace cress {

// Stand-alone Thomas solve over a box of columns for the s_gaussel shim
// (gaussel.f90:169-262, impopt=1), one lane per column.
__global__ void k_gaussel(Grid g, int istr, int iend, int jstr, int jend, int kstr, int kend,
                          const float *rr, float *ss, const float *tt, float *ff) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nci = iend - istr + 1;
  const int ncol = nci * (jend - jstr + 1);
  if (t >= ncol) return;
  const int i = istr + t % nci, j = jstr + t / nci;
  const int kem1 = kend - 1;
  if (kstr == kend) {
    AT(ff, i, j, kstr) = AT(ff, i, j, kstr) / AT(ss, i, j, kstr);
  } else {
    const float aa = 1.0f / AT(ss, i, j, kstr);
    AT(ss, i, j, kstr) = AT(tt, i, j, kstr) * aa;
    AT(ff, i, j, kstr) = AT(ff, i, j, kstr) * aa;
  }
  if (kstr + 2 <= kend) {
    for (int k = kstr + 1; k <= kend - 1; k++) {
      const float aa = 1.0f / (AT(ss, i, j, k) - AT(rr, i, j, k) * AT(ss, i, j, k - 1));
      AT(ss, i, j, k) = AT(tt, i, j, k) * aa;
      AT(ff, i, j, k) = (AT(ff, i, j, k) - AT(rr, i, j, k) * AT(ff, i, j, k - 1)) * aa;
    }
  }
  if (kstr + 1 <= kend) {
    const float aa = 1.0f / (AT(ss, i, j, kend) - AT(rr, i, j, kend) * AT(ss, i, j, kem1));
    AT(ff, i, j, kend) = (AT(ff, i, j, kend) - AT(rr, i, j, kend) * AT(ff, i, j, kem1)) * aa;
    for (int k = kend - 1; k >= kstr; k--) AT(ff, i, j, k) = AT(ff, i, j, k) - AT(ss, i, j, k) * AT(ff, i, j, k + 1);
  }
}

// advs alone (advs.f90:205-445) for the s_advs shim: sfrc over the interior.
__global__ void k_advs_only(Grid g, Opts o, Consts c, ScalarIn in) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > g.ni - 2 || j > g.nj - 2) return;
  AT(in.frc[0], i, j, k) = adv_s(g, o, c, in.s[0], in.rstxu, in.rstxv, in.rstxwc, i, j, k);
}

// s_advuvw alone (advuvw.f90:75-912): ufrc, vfrc, wfrc plus the advection
// of u (i=2..ni-1, j=2..nj-2, k=2..nk-2), v (i=2..ni-2, j=2..nj-1) and w
// (i=2..ni-2, j=2..nj-2, k=2..nk-1), the terms and order of the forcing
// kernels.
__global__ void k_advuvw_only(Grid g, Opts o, Consts c, const float *u, const float *v, const float *w,
                              const float *rxu, const float *rxv, const float *rxw, float *ufrc, float *vfrc,
                              float *wfrc) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > g.ni - 1 || j > g.nj - 1) return;
  if (k <= g.nk - 2 && j <= g.nj - 2) AT(ufrc, i, j, k) = adv_u(AT(ufrc, i, j, k), g, o, c, u, rxu, rxv, rxw, i, j, k);
  if (k <= g.nk - 2 && i <= g.ni - 2) AT(vfrc, i, j, k) = adv_v(AT(vfrc, i, j, k), g, o, c, v, rxu, rxv, rxw, i, j, k);
  if (i <= g.ni - 2 && j <= g.nj - 2) AT(wfrc, i, j, k) = adv_w(AT(wfrc, i, j, k), g, o, c, w, rxu, rxv, rxw, i, j, k);
}

// s_bc8u, s_bc8v (bc8u.f90:132-202, bc8v.f90), s_bc8w (bc8w.f90:130-184) on
// one rank, every side an outer boundary: the staggered point on the
// boundary takes the value two points in (rigid, option 2) or one point in
// (options >= 3); periodic sides (-1) are left alone.  kind 0 x (u points,
// planes i=1 and ni), 1 y (v points), 2 z (w points, levels 1 and nk).
// One thread per point of the face pair.
__global__ void k_bc8(Grid g, int kind, int lo, int hi, int kmax, float *var) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (kind == 2) {
    const int i = t % (g.ni + 2), j = t / (g.ni + 2);
    if (j > g.nj + 1) return;
    if (m == 0 && lo >= 2) AT(var, i, j, 1) = AT(var, i, j, lo == 2 ? 3 : 2);
    if (m == 1 && hi >= 2) AT(var, i, j, g.nk) = AT(var, i, j, hi == 2 ? g.nk - 2 : g.nk - 1);
    return;
  }
  const int n = kind == 0 ? g.nj + 2 : g.ni + 2;
  const int a = t % n, k = 1 + t / n;
  if (k > kmax) return;
  if (kind == 0) {
    if (m == 0 && lo >= 2) AT(var, 1, a, k) = AT(var, lo == 2 ? 3 : 2, a, k);
    if (m == 1 && hi >= 2) AT(var, g.ni, a, k) = AT(var, hi == 2 ? g.ni - 2 : g.ni - 1, a, k);
  } else {
    if (m == 0 && lo >= 2) AT(var, a, 1, k) = AT(var, a, lo == 2 ? 3 : 2, k);
    if (m == 1 && hi >= 2) AT(var, a, g.nj, k) = AT(var, a, hi == 2 ? g.nj - 2 : g.nj - 1, k);
  }
}

// The forcing kernels with and without the smoothing compiled in.
template __global__ void k_ac_ptuv<false, 0>(Tiles, Grid, Opts, Consts, AcPtUv);
template __global__ void k_ac_ptuv<true, 0>(Tiles, Grid, Opts, Consts, AcPtUv);
template __global__ void k_ac_ptuv<false, 1>(Tiles, Grid, Opts, Consts, AcPtUv);
template __global__ void k_ac_ptuv<true, 1>(Tiles, Grid, Opts, Consts, AcPtUv);
template __global__ void k_ac_col<false>(Tiles, Grid, Opts, Consts, AcCol);
template __global__ void k_ac_col<true>(Tiles, Grid, Opts, Consts, AcCol);
#define CRESS_INST_U(U)                                                                                      \
  template __global__ void k_phy2cnt<U>(Tiles, Grid, Opts, const float *, const float *, float *);           \
  template __global__ void k_prep<U>(Tiles, Grid, Opts, const float *, const float *, const float *,         \
                                     const float *, const float *, const float *, const float *, float *,   \
                                     float *, float *, float *);                                             \
  template __global__ void k_force_p<U>(Tiles, Grid, Opts, Consts, const float *, const float *,             \
                                        const float *, const float *, const float *, const float *,          \
                                        const float *, const float *, const float *, const float *, float *);\
  template __global__ void k_force_uv<false, U>(Tiles, Grid, Opts, Consts, ForceIn, float *, float *);       \
  template __global__ void k_force_uv<true, U>(Tiles, Grid, Opts, Consts, ForceIn, float *, float *);        \
  template __global__ void k_force_w<false, U>(Tiles, Grid, Opts, Consts, ForceIn, float *);                 \
  template __global__ void k_force_w<true, U>(Tiles, Grid, Opts, Consts, ForceIn, float *);                  \
  template __global__ void k_force_s<false, U>(Tiles, Grid, Opts, Consts, ScalarIn, int);                    \
  template __global__ void k_force_s<true, U>(Tiles, Grid, Opts, Consts, ScalarIn, int);
CRESS_INST_U(false)
CRESS_INST_U(true)

}  // namespace cress
