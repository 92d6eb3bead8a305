// Kernels of the TKE-closure coefficients (turbdrv.f90:245-276, tubopt=2).
// Each writes the index ranges the reference writes; the arithmetic keeps
// Fortran's left-to-right order term by term, so with -ffp-contract=off
// the tensor, its square and dry / cloud-free N^2 are the reference's
// bits.  exp and log are binary32 with float-float internals (xmath.h; N^2
// in cloudy air, the length scale ln0 = (ds^3 jcb)^(1/3) of the eddy
// viscosity, taken as exp(log(x)/3) as the reference writes it): within
// 1e-5 of the reference's FP32 libm results.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"
#include "dyn.h"
#include "turb.h"
#include "xmath.h"

namespace cress {

#define AT(a, i, j, k) a[g.at(i, j, k)]

// exp and log in binary32 (xmath.h: the correctly rounded result but for
// arguments within ~2^-40 of a rounding boundary); the reference's FP32
// build uses glibc's, which agrees on all but ~1e-4 of arguments.
__device__ __forceinline__ float expd(float x) { return xexpf(x); }
__device__ __forceinline__ float logd(float x) { return xlogf(x); }

// comphy.f90:49-130, commath.f90:51,177
#define TB_G 9.8f
#define TB_CP 1004.f
#define TB_CW 4.218e3f
#define TB_CI 2.106e3f
#define TB_RD 287.f
#define TB_P0 1.e5f
#define TB_T0 273.16f
#define TB_TLOW 233.16f
#define TB_LV0 2.50078e6f
#define TB_LF0 3.34e5f
#define TB_EPSVA .622f
#define TB_EPSAV 1.6077f
#define TB_EPS 1.e-20f

// The moist potential temperature and the factor a of bruntv.f90:276-300
// at one cell (cloud physics on).
__device__ __forceinline__ void moist_pt_a(const Grid &g, const TurbArgs &a, int i, int j, int k, float &pt,
                                           float &fa) {
  const float rddvcp = TB_RD / TB_CP, cwmci = TB_CW - TB_CI, p0iv = 1.f / TB_P0;
  const float qv = AT(a.qv, i, j, k);
  pt = AT(a.ptbr, i, j, k) + AT(a.ptp, i, j, k);
  const float t = pt * expd(rddvcp * logd(p0iv * (AT(a.pbr, i, j, k) + AT(a.pp, i, j, k))));
  float ak = TB_LV0 * expd((.167f + 3.67e-4f * t) * logd(TB_T0 / t));
  if (t <= TB_TLOW) ak = ak + (TB_LF0 + cwmci * (t - TB_T0));
  const float lhcpt = ak / (TB_CP * t);
  pt = pt * expd(lhcpt * qv);
  ak = ak * qv / (TB_RD * t);
  fa = (1.f + ak) / (1.f + TB_EPSVA * lhcpt * ak);
}

__device__ __forceinline__ float ptv_at(const Grid &g, const TurbArgs &a, int i, int j, int k) {
  const float pt = AT(a.ptbr, i, j, k) + AT(a.ptp, i, j, k);
  if (!a.moist) return pt;
  const float qv = AT(a.qv, i, j, k);
  return pt * (1.f + TB_EPSAV * qv) / (1.f + qv);
}

// s_bruntv: N^2 at w points over i 1..ni-1, j 1..nj-1, k 2..nk-1, and its
// copies to k=1 and k=nk (bruntv.f90:374-384).
__global__ void k_bruntv(Grid g, TurbArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 1) return;
  const int kk = k == 1 ? 2 : k == g.nk ? g.nk - 1 : k;
  const float dziv = 1.f / a.dz;
  const float gdzv = TB_G * dziv, gdzv05 = .5f * TB_G * dziv;
  const float den = AT(a.jcb8w, i, j, kk) * (AT(a.ptbr, i, j, kk - 1) + AT(a.ptbr, i, j, kk));
  float n;
  if (a.moist && a.cphopt != 0 && AT(a.qall, i, j, kk) > a.thresq) {
    float p0, a0, p1, a1;
    moist_pt_a(g, a, i, j, kk - 1, p0, a0);
    moist_pt_a(g, a, i, j, kk, p1, a1);
    n = gdzv05 *
        ((AT(a.qall, i, j, kk - 1) - AT(a.qall, i, j, kk)) +
         (p1 - p0) * (a0 + a1) / (AT(a.ptbr, i, j, kk - 1) + AT(a.ptbr, i, j, kk))) /
        AT(a.jcb8w, i, j, kk);
  } else {
    n = gdzv * (ptv_at(g, a, i, j, kk) - ptv_at(g, a, i, j, kk - 1)) / den;
  }
  AT(a.nsq8w, i, j, k) = n;
}

// s_defomten, trnopt=0 and mfcopt=0 (defomten.f90:253-1090): every
// component at one (i,j,k), over the ranges the reference writes.
__global__ void k_defomten(Grid g, TurbArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni + 1) return;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const float dxiv = 1.f / a.dx, dyiv = 1.f / a.dy, dziv = 1.f / a.dz;
#define UJ(i, j, k) (AT(a.u, i, j, k) * AT(a.jcb8u, i, j, k))
#define VJ(i, j, k) (AT(a.v, i, j, k) * AT(a.jcb8v, i, j, k))
#define WJ(i, j, k) (AT(a.w, i, j, k) * AT(a.jcb8w, i, j, k))
  if (k <= nk - 1) {
    if (j >= 1 + a.js && j <= nj - a.jn && i >= 1 + a.iw && i <= ni - a.ie) {
      const float t3 = 4.f / ((AT(a.jcb, i - 1, j - 1, k) + AT(a.jcb, i, j, k)) +
                              (AT(a.jcb, i - 1, j, k) + AT(a.jcb, i, j - 1, k)));
      AT(a.s12, i, j, k) = t3 * ((UJ(i, j, k) - UJ(i, j - 1, k)) * dyiv + (VJ(i, j, k) - VJ(i - 1, j, k)) * dxiv);
    }
    if (j >= 1 && j <= nj - 1 && i >= 1 && i <= ni - 1) {
      const float jcbiv2 = 2.f / AT(a.jcb, i, j, k);
      AT(a.s11, i, j, k) = jcbiv2 * ((UJ(i + 1, j, k) - UJ(i, j, k)) * dxiv);
      AT(a.s22, i, j, k) = jcbiv2 * ((VJ(i, j + 1, k) - VJ(i, j, k)) * dyiv);
      AT(a.s33, i, j, k) = jcbiv2 * (AT(a.w, i, j, k + 1) - AT(a.w, i, j, k)) * dziv;
    }
  }
  if (k >= 2 && k <= nk - 1) {
    if (j >= 1 && j <= nj - 1 && i >= 1 + a.iw && i <= ni - a.ie) {
      const float t3 = 2.f / (AT(a.jcb8w, i - 1, j, k) + AT(a.jcb8w, i, j, k));
      const float s = ((WJ(i, j, k) - WJ(i - 1, j, k)) * dxiv + (AT(a.u, i, j, k) - AT(a.u, i, j, k - 1)) * dziv) * t3;
      AT(a.s13, i, j, k) = s;
      AT(a.s31, i, j, k) = s;
    }
    if (j >= 1 + a.js && j <= nj - a.jn && i >= 1 && i <= ni - 1) {
      const float t3 = 2.f / (AT(a.jcb8w, i, j - 1, k) + AT(a.jcb8w, i, j, k));
      const float s = ((WJ(i, j, k) - WJ(i, j - 1, k)) * dyiv + (AT(a.v, i, j, k) - AT(a.v, i, j, k - 1)) * dziv) * t3;
      AT(a.s23, i, j, k) = s;
      AT(a.s32, i, j, k) = s;
    }
  }
#undef UJ
#undef VJ
#undef WJ
}

// s_bcten (bcten.f90) on s31 and s32: level 1 and nk from 3 and nk-2.
__global__ void k_bcten(Grid g, int bbc, int tbc, float *s31, float *s32) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.nxy) return;
  float *t = blockIdx.y ? s32 : s31;
  const long lo = g.nxy * 0, l3 = g.nxy * 2, hi = g.nxy * (g.nk - 1), h2 = g.nxy * (g.nk - 3);
  if (bbc == 2) t[lo + p] = -t[l3 + p];
  else if (bbc >= 3) t[lo + p] = t[l3 + p];
  if (tbc == 2) t[hi + p] = -t[h2 + p];
  else if (tbc >= 3) t[hi + p] = t[h2 + p];
}

// s_defomssq (defomssq.f90:95-120) and s_eddyvis (eddyvis.f90:454-486,
// tubopt=2, sfcopt=0, isoopt=1, mfcopt=0) at one interior cell.
__global__ void k_eddyvis(Grid g, TurbArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 1) return;
  const float a12 = (AT(a.s12, i, j, k) + AT(a.s12, i + 1, j + 1, k)) + (AT(a.s12, i + 1, j, k) + AT(a.s12, i, j + 1, k));
  const float a31 = (AT(a.s31, i, j, k) + AT(a.s31, i + 1, j, k + 1)) + (AT(a.s31, i, j, k + 1) + AT(a.s31, i + 1, j, k));
  const float a32 = (AT(a.s32, i, j, k) + AT(a.s32, i, j + 1, k + 1)) + (AT(a.s32, i, j + 1, k) + AT(a.s32, i, j, k + 1));
  const float s11 = AT(a.s11, i, j, k), s22 = AT(a.s22, i, j, k), s33 = AT(a.s33, i, j, k);
  const float ssq = .5f * (s33 * s33 + (s11 * s11 + s22 * s22)) + .0625f * (a12 * a12 + (a31 * a31 + a32 * a32));
  AT(a.ssq, i, j, k) = ssq;
  const float ds3 = a.dx * a.dy * a.dz;
  const float ckm = .1f, ckmin = 1.e-6f, ckmax = .125f / a.dtb;
  const float ln0 = expd((1.f / 3.f) * logd(ds3 * AT(a.jcb, i, j, k)));
  const float ln02 = ln0 * ln0;
  const float nsq = AT(a.nsq8w, i, j, k) + AT(a.nsq8w, i, j, k + 1);
  const float tke = AT(a.tke, i, j, k);
  float ln;
  if (nsq < 0.f) {
    ln = ln0;
  } else {
    float m = .76f * sqrtf(tke / (nsq + TB_EPS));
    if (!(m < ln0)) m = ln0;
    ln = .1f * ln0;
    if (!(ln > m)) ln = m;
  }
  const float priv = 1.f + 2.f * ln / ln0;
  float kv = ckm * ln * sqrtf(tke);
  if (priv * nsq < ssq) {
    const float lo = ckmin * ln02;
    if (lo > kv) kv = lo;
  }
  const float hi = ckmax * ln02;
  kv = AT(a.rbr, i, j, k) * (hi < kv ? hi : kv);
  AT(a.priv, i, j, k) = priv;
  AT(a.rkv, i, j, k) = kv;
  AT(a.rkh, i, j, k) = kv;
}

void turb_coef(const Grid &g, const TurbArgs &a) {
  const dim3 blk(64);
  hipLaunchKernelGGL(k_bruntv, dim3((g.ni - 1 + 63) / 64, g.nj - 1, g.nk), blk, 0, 0, g, a);
  hipLaunchKernelGGL(k_defomten, dim3((g.ni + 2 + 63) / 64, g.nj + 2, g.nk), blk, 0, 0, g, a);
  // defomten.f90:1095-1102
  const unsigned px = (unsigned)(((long)(g.nj + 2) * g.nk + 63) / 64), py = (unsigned)(((long)(g.ni + 2) * g.nk + 63) / 64);
  hipLaunchKernelGGL(k_bc8, dim3(px, 2), blk, 0, 0, g, 0, a.wbc, a.ebc, g.nk, a.s12);
  hipLaunchKernelGGL(k_bc8, dim3(py, 2), blk, 0, 0, g, 1, a.sbc, a.nbc, g.nk, a.s12);
  hipLaunchKernelGGL(k_bc8, dim3(px, 2), blk, 0, 0, g, 0, a.wbc, a.ebc, g.nk, a.s31);
  hipLaunchKernelGGL(k_bc8, dim3(py, 2), blk, 0, 0, g, 1, a.sbc, a.nbc, g.nk, a.s32);
  hipLaunchKernelGGL(k_bcten, dim3((unsigned)((g.nxy + 63) / 64), 2), blk, 0, 0, g, a.bbc, a.tbc, a.s31, a.s32);
  hipLaunchKernelGGL(k_eddyvis, dim3((g.ni - 1 + 63) / 64, g.nj - 1, g.nk - 1), blk, 0, 0, g, a);
}

// s_strsten (strsten.f90:150-210), sfcopt=0: the tensor times the eddy
// viscosities, in place (each point reads only its own component).
__global__ void k_strsten(Grid g, TurbArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 1) return;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const float *h = a.rkh, *v = a.rkv;
  AT(a.s11, i, j, k) = AT(h, i, j, k) * AT(a.s11, i, j, k);
  AT(a.s22, i, j, k) = AT(h, i, j, k) * AT(a.s22, i, j, k);
  AT(a.s33, i, j, k) = AT(v, i, j, k) * AT(a.s33, i, j, k);
  if (i >= 2 && j >= 2)
    AT(a.s12, i, j, k) = .25f * AT(a.s12, i, j, k) *
                         ((AT(h, i - 1, j - 1, k) + AT(h, i, j, k)) + (AT(h, i - 1, j, k) + AT(h, i, j - 1, k)));
  if (k < 2 || i < 2 || j < 2) return;
  if (j <= nj - 2) {
    AT(a.s13, i, j, k) = .25f * AT(a.s13, i, j, k) *
                         ((AT(v, i - 1, j, k - 1) + AT(v, i, j, k)) + (AT(v, i - 1, j, k) + AT(v, i, j, k - 1)));
    AT(a.s31, i, j, k) = .25f * AT(a.s31, i, j, k) *
                         ((AT(h, i - 1, j, k - 1) + AT(h, i, j, k)) + (AT(h, i - 1, j, k) + AT(h, i, j, k - 1)));
  }
  if (i <= ni - 2) {
    AT(a.s23, i, j, k) = .25f * AT(a.s23, i, j, k) *
                         ((AT(v, i, j - 1, k - 1) + AT(v, i, j, k)) + (AT(v, i, j - 1, k) + AT(v, i, j, k - 1)));
    AT(a.s32, i, j, k) = .25f * AT(a.s32, i, j, k) *
                         ((AT(h, i, j - 1, k - 1) + AT(h, i, j, k)) + (AT(h, i, j - 1, k) + AT(h, i, j, k - 1)));
  }
  (void)nk;
}

// s_turbuvw (turbuvw.f90:288-352, 795-859, 1300-1330), trnopt=0, mfcopt=0,
// the leapfrog branch: the stress divergence overwrites the forcings.  The
// reference's rescaled t11 / t22 and its tmp1 products are formed where
// used (the same single products).
__global__ void k_turbuvw(Grid g, TurbArgs a) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y, k = 2 + blockIdx.z;
  if (i > g.ni - 1) return;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const float dxiv = 1.f / a.dx, dyiv = 1.f / a.dy, dziv = 1.f / a.dz;
  const float dxiv05 = .5f * dxiv, dyiv05 = .5f * dyiv;
#define TU(i, j, k) ((AT(a.jcb8u, i, j - 1, k) + AT(a.jcb8u, i, j, k)) * AT(a.s12, i, j, k))
#define TV(i, j, k) ((AT(a.jcb8v, i - 1, j, k) + AT(a.jcb8v, i, j, k)) * AT(a.s12, i, j, k))
#define TW1(i, j, k) ((AT(a.jcb8u, i, j, k - 1) + AT(a.jcb8u, i, j, k)) * AT(a.s31, i, j, k))
#define TW2(i, j, k) ((AT(a.jcb8v, i, j, k - 1) + AT(a.jcb8v, i, j, k)) * AT(a.s32, i, j, k))
  if (k <= nk - 2 && j <= nj - 2) {
    const float t0 = AT(a.jcb, i, j, k) * AT(a.s11, i, j, k), tm = AT(a.jcb, i - 1, j, k) * AT(a.s11, i - 1, j, k);
    AT(a.ufrc, i, j, k) = ((t0 - tm) * dxiv + (TU(i, j + 1, k) - TU(i, j, k)) * dyiv05) +
                          (AT(a.s13, i, j, k + 1) - AT(a.s13, i, j, k)) * dziv;
  }
  if (k <= nk - 2 && i <= ni - 2) {
    const float t0 = AT(a.jcb, i, j, k) * AT(a.s22, i, j, k), tm = AT(a.jcb, i, j - 1, k) * AT(a.s22, i, j - 1, k);
    AT(a.vfrc, i, j, k) = ((t0 - tm) * dyiv + (TV(i + 1, j, k) - TV(i, j, k)) * dxiv05) +
                          (AT(a.s23, i, j, k + 1) - AT(a.s23, i, j, k)) * dziv;
  }
  if (j <= nj - 2 && i <= ni - 2)
    AT(a.wfrc, i, j, k) = (AT(a.s33, i, j, k) - AT(a.s33, i, j, k - 1)) * dziv +
                          ((TW1(i + 1, j, k) - TW1(i, j, k)) * dxiv05 + (TW2(i, j + 1, k) - TW2(i, j, k)) * dyiv05);
#undef TU
#undef TV
#undef TW1
#undef TW2
}

void turb_drv(const Grid &g, const TurbArgs &a) {
  turb_coef(g, a);
  const dim3 blk(64);
  hipLaunchKernelGGL(k_strsten, dim3((g.ni - 1 + 63) / 64, g.nj - 1, g.nk - 1), blk, 0, 0, g, a);
  hipLaunchKernelGGL(k_bcten, dim3((unsigned)((g.nxy + 63) / 64), 2), blk, 0, 0, g, a.bbc, a.tbc, a.s31, a.s32);
  hipLaunchKernelGGL(k_turbuvw, dim3((g.ni - 2 + 63) / 64, g.nj - 2, g.nk - 2), blk, 0, 0, g, a);
}

// s_turbflx (turbflx.f90:180-262), trnopt=0, sfcopt=0: the turbulent
// fluxes of one scalar at u, v and w points over the reference's ranges.
__global__ void k_turbflx(Grid g, float dx, float dy, float dz, const float *jcb, const float *s,
                          const float *kh8u, const float *kh8v, const float *kv8w, float *h1, float *h2, float *h3) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 1) return;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const float dxiv = 1.f / dx, dyiv = 1.f / dy, dziv = 1.f / dz;
  const float dxv05 = .5f * dxiv, dyv05 = .5f * dyiv;
  const float js = AT(s, i, j, k) * AT(jcb, i, j, k);
  if (j <= nj - 2) AT(h1, i, j, k) = AT(kh8u, i, j, k) * (js - AT(s, i - 1, j, k) * AT(jcb, i - 1, j, k)) * dxv05;
  if (i <= ni - 2) AT(h2, i, j, k) = AT(kh8v, i, j, k) * (js - AT(s, i, j - 1, k) * AT(jcb, i, j - 1, k)) * dyv05;
  if (k >= 2 && i <= ni - 2 && j <= nj - 2)
    AT(h3, i, j, k) = AT(kv8w, i, j, k) * (AT(s, i, j, k) - AT(s, i, j, k - 1)) * dziv;
  (void)nk;
}

void turb_flx(const Grid &g, float dx, float dy, float dz, const float *jcb, const float *s, const float *kh8u,
              const float *kh8v, const float *kv8w, float *h1, float *h2, float *h3) {
  hipLaunchKernelGGL(k_turbflx, dim3((g.ni - 2 + 63) / 64, g.nj - 2, g.nk - 1), dim3(64), 0, 0, g, dx, dy, dz, jcb,
                     s, kh8u, kh8v, kv8w, h1, h2, h3);
}

// s_turbs (turbs.f90:189-212), trnopt=0, mfcopt=0: the divergence of the
// fluxes added to the scalar's forcing over i,j 2..n-2, k 2..nk-2.
__global__ void k_turbs(Grid g, float dx, float dy, float dz, const float *jcb8u, const float *jcb8v,
                        const float *h1, const float *h2, const float *h3, float *sfrc) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y, k = 2 + blockIdx.z;
  if (i > g.ni - 2) return;
  const float dxiv = 1.f / dx, dyiv = 1.f / dy, dziv = 1.f / dz;
  AT(sfrc, i, j, k) = AT(sfrc, i, j, k) +
                      ((AT(h3, i, j, k + 1) - AT(h3, i, j, k)) * dziv +
                       ((AT(jcb8u, i + 1, j, k) * AT(h1, i + 1, j, k) - AT(jcb8u, i, j, k) * AT(h1, i, j, k)) * dxiv +
                        (AT(jcb8v, i, j + 1, k) * AT(h2, i, j + 1, k) - AT(jcb8v, i, j, k) * AT(h2, i, j, k)) * dyiv));
}

void turb_s(const Grid &g, float dx, float dy, float dz, const float *jcb8u, const float *jcb8v, const float *h1,
            const float *h2, const float *h3, float *sfrc) {
  hipLaunchKernelGGL(k_turbs, dim3((g.ni - 3 + 63) / 64, g.nj - 3, g.nk - 3), dim3(64), 0, 0, g, dx, dy, dz, jcb8u,
                     jcb8v, h1, h2, h3, sfrc);
}

// s_kh8uv (kh8uv.f90:214-232), mfcopt=0: rkh summed onto u and v points.
__global__ void k_kh8uv(Grid g, const float *rkh, float *rkh8u, float *rkh8v) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 1) return;
  if (i >= 2) AT(rkh8u, i, j, k) = AT(rkh, i - 1, j, k) + AT(rkh, i, j, k);
  if (j >= 2) AT(rkh8v, i, j, k) = AT(rkh, i, j - 1, k) + AT(rkh, i, j, k);
}

void turb_kh8uv(const Grid &g, const float *rkh, float *rkh8u, float *rkh8v) {
  hipLaunchKernelGGL(k_kh8uv, dim3((g.ni - 1 + 63) / 64, g.nj - 1, g.nk - 1), dim3(64), 0, 0, g, rkh, rkh8u, rkh8v);
}

// s_sheartke (sheartke.f90:93-101) and s_disptke (disptke.f90:180-195,
// isoopt=1, mfcopt=0) over i,j 2..n-2, k 2..nk-2, then s_eddyvisj
// (eddyvisj.f90:103-112, mfcopt=0) over i,j 1..n-1, k 1..nk-1; each cell
// reads and writes only itself.  The length scale is (dx dy dz jcb / 8)^(1/3)
// taken as exp(log(x)/3), as the reference writes it.
__global__ void k_tke_src(Grid g, TkeArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 1) return;
  const float jcb = AT(a.jcb, i, j, k), rkv = AT(a.rkv, i, j, k);
  if (i >= 2 && i <= g.ni - 2 && j >= 2 && j <= g.nj - 2 && k >= 2 && k <= g.nk - 2) {
    float f = AT(a.tkefrc, i, j, k) + jcb * rkv * AT(a.ssq, i, j, k);
    const float ds308 = .125f * a.dx * a.dy * a.dz;
    const float pv = AT(a.priv, i, j, k), e = AT(a.tkep, i, j, k);
    const float ln = (pv - 1.f) * expd((1.f / 3.f) * logd(ds308 * jcb)) + TB_EPS;
    AT(a.tkefrc, i, j, k) = f - (.37f * pv - .18f) * AT(a.rst, i, j, k) * e * sqrtf(e) / ln;
  }
  const float jcbiv = 1.f / jcb;
  AT(a.rkh, i, j, k) = jcbiv * AT(a.rkh, i, j, k);
  AT(a.rkv, i, j, k) = jcbiv * rkv;
}

// s_tkeflx (tkeflx.f90:181-330, trnopt=0, mfcopt=0) and s_turbtke
// (turbtke.f90:191-213) fused: each cell forms the six fluxes on its faces
// from rkh, rkv (already divided by jcb) and the past TKE with the
// reference's operations, so the divergence is the reference's bits; then
// s_vsps0 (vsps0.f90:114-127) from level ksp0(2)-1.  k runs from 1 so that
// the sponge reaches level 1 when ksp0(2) = 2, as the reference's does.
__global__ void k_tke_flx(Grid g, TkeArgs a) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 2 + blockIdx.y, k = 1 + blockIdx.z;
  if (i > g.ni - 2) return;
  const bool sponge = a.vsp && k >= a.ksp0 - 1;
  if (k == 1 && !sponge) return;
  float f = AT(a.tkefrc, i, j, k);
  if (k >= 2) {
    const float dxiv = 1.f / a.dx, dyiv = 1.f / a.dy, dziv = 1.f / a.dz;
    const float dxv5 = .5f * dxiv, dyv5 = .5f * dyiv, dzv5 = .5f * dziv;
    const float dxiv2 = 2.f * dxiv, dyiv2 = 2.f * dyiv, dziv2 = 2.f * dziv;
    auto jt = [&](int ii, int jj) { return AT(a.tkep, ii, jj, k) * AT(a.jcb, ii, jj, k); };
    auto h1 = [&](int ii) { return (AT(a.rkh, ii - 1, j, k) + AT(a.rkh, ii, j, k)) * (jt(ii, j) - jt(ii - 1, j)) * dxv5; };
    auto h2 = [&](int jj) { return (AT(a.rkh, i, jj - 1, k) + AT(a.rkh, i, jj, k)) * (jt(i, jj) - jt(i, jj - 1)) * dyv5; };
    auto h3 = [&](int kk) {
      return (AT(a.rkv, i, j, kk - 1) + AT(a.rkv, i, j, kk)) * (AT(a.tkep, i, j, kk) - AT(a.tkep, i, j, kk - 1)) * dzv5;
    };
    f = f + ((h3(k + 1) - h3(k)) * dziv2 +
             ((AT(a.jcb8u, i + 1, j, k) * h1(i + 1) - AT(a.jcb8u, i, j, k) * h1(i)) * dxiv2 +
              (AT(a.jcb8v, i, j + 1, k) * h2(j + 1) - AT(a.jcb8v, i, j, k) * h2(j)) * dyiv2));
  }
  if (sponge) {
    // rbct(i,j,k,2) and rbct(i,j,k+1,2)
    const size_t nij = (size_t)g.ni * g.nj;
    const size_t o = (size_t)(i - 1) + (size_t)g.ni * (j - 1) + nij * ((size_t)(k - 1) + (size_t)g.nk);
    f = f - .5f * (a.rbct[o] + a.rbct[o + nij]) * AT(a.rst, i, j, k) * AT(a.tkep, i, j, k);
  }
  AT(a.tkefrc, i, j, k) = f;
}

// s_eddydif (eddydif.f90:266-330, tubopt>=2, isoopt=1) and s_buoytke
// (buoytke.f90:96-115) down one column per thread, i,j 1..n-1: rkv8s =
// priv rkv, rkh = rkv8s, rkv = the mean of rkv8s over the two levels
// (k >= 2), rkv8s halved and times jcb; the buoyancy flux at w levels is
// carried from one level to the next, and the TKE forcing of level k-1
// takes the two around it.  Each column reads and writes only itself.
__global__ void k_tke_dif(Grid g, TkeArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y;
  if (i > g.ni - 1) return;
  const bool inner = i >= 2 && i <= g.ni - 2 && j >= 2 && j <= g.nj - 2;
  float sp = 0.f, r8p = 0.f, bp = 0.f;
  for (int k = 1; k <= g.nk - 1; k++) {
    const float s = AT(a.priv, i, j, k) * AT(a.rkv, i, j, k);
    AT(a.rkh, i, j, k) = s;
    if (k >= 2) AT(a.rkv, i, j, k) = .5f * (sp + s);
    const float r8 = .5f * AT(a.jcb, i, j, k) * s;
    AT(a.rkv8s, i, j, k) = r8;
    if (inner && k >= 2) {
      const float b = AT(a.jcb8w, i, j, k) * AT(a.nsq8w, i, j, k) * (r8p + r8);
      if (k >= 3) AT(a.tkefrc, i, j, k - 1) = AT(a.tkefrc, i, j, k - 1) - (bp + b);
      bp = b;
    }
    sp = s;
    r8p = r8;
  }
}

void turb_tke(const Grid &g, const TkeArgs &a) {
  const dim3 blk(64);
  hipLaunchKernelGGL(k_tke_src, dim3((g.ni - 1 + 63) / 64, g.nj - 1, g.nk - 1), blk, 0, 0, g, a);
  hipLaunchKernelGGL(k_tke_flx, dim3((g.ni - 3 + 63) / 64, g.nj - 3, g.nk - 2), blk, 0, 0, g, a);
  hipLaunchKernelGGL(k_tke_dif, dim3((g.ni - 1 + 63) / 64, g.nj - 1), blk, 0, 0, g, a);
}

}  // namespace cress
