// Cold-rain bulk microphysics (cphopt=2, haiopt=0) for gfx950.
//
// s_cloudphy -> s_coldblk (cloudphy.f90:277-338, coldblk.f90:321-481) as
//   k_micro_src : one lane per cell, all of getvdens, getexner, diagnw,
//                 diagnsg, adjstnci, termblk and the srcblk chain (setblk,
//                 nuc1stv, nuc1stc, collect, prodctwg, distrpg, aggregat,
//                 melting, depsit, convers, nuc2nd, freezing, shedding,
//                 more0q, newblk) in registers, then adjstnci; it writes
//                 the three precipitation fall speeds and folds the
//                 sedimentation CFL minimum (fallblk.f90:311-437) into one
//                 device word;
//   k_micro_fall: one lane per column, fallblk's npstp upwind sub-steps
//                 (upwqp.f90, upwnp.f90) with the surface precipitation;
// and s_satadjst (satadjst.f90:154-187) as k_satadjst, one lane per cell.
// Rates follow the reference's formulas and operation order; options fixed:
// flqcqi_opt=1, evapor_opt=1, nuc2nd_opt=1, advopt<=3.
#include "micro.h"
#include "xmath.h"

namespace cress {

namespace mc {
// comphy.f90, commath.f90, temparam.f90 (UserModAdd overlay)
constexpr float g = 9.8f, cp = 1004.0f, cw = 4.218e3f, ci = 2.106e3f, rd = 287.0f, rv = 461.0f;
constexpr float epsva = 0.622f, es0 = 610.78f, r0 = 1.225f, p0 = 1.e5f, t0 = 273.16f;
constexpr float tlow = 233.16f, t0cel = 0.0f, t20 = 293.16f, t23 = 296.16f, p20 = 101325.0f;
constexpr float lmb20 = 6.6e-8f, diasl = 6.e-7f, boltz = 1.38066e-23f, sc = 0.6f;
constexpr float lv0 = 2.50078e6f, lf0 = 3.34e5f, kp0 = 10.79f, kpa = 2.4e-1f, nu0 = 1.328e-5f;
constexpr float dv0 = 2.23e-5f, xl = 0.25f, nr0 = 8.e6f, ns0 = 1.8e6f, ng0 = 1.1e6f;
constexpr float aur = 842.0f, aui = 770.0f, aus = 27.0f, aug = 132.0f;
constexpr float bur = 0.8f, bus = 0.5f, bug = 0.64f, gur = 0.5f, gus = 0.5f, gug = 0.5f;
constexpr float rhow = 1.e3f, rhoi = 2.e2f, rhos = 1.e2f, rhog = 4.e2f;
constexpr float ers = 1.0f, erg = 1.0f, eir = 1.0f, eis = 0.1f, eig = 0.1f, esr = 1.0f, esg = 0.001f;
constexpr float ecc = 0.55f, eii = 0.1f, ess = 0.1f;
constexpr float mr0 = 2.7e-10f, mi0 = 1.e-12f, ms0 = 1.76e-10f, mg0 = 7.e-10f;
constexpr float mrmax = 2.68e-4f, mimax = 1.4e-10f, msmax = 6.e-4f, mgmax = 8.e-3f;
constexpr float nclcst = 1.e8f, diaqcm = 2.e-5f, eps = 1.e-20f, cc = 3.141592f, ln10 = 2.30259f;
constexpr float gf1bur = 0.931384f, gf1bus = 0.886227f, gf1bug = 0.898642f;
constexpr float gf3bur = 4.69417f, gf3bus = 3.32335f, gf3bug = 3.89076f;
constexpr float gf4bur = 17.8379f, gf4bus = 11.6314f, gf4bug = 14.1624f;
constexpr float gf5bur = 1.82736f, gf5bus = 1.60836f, gf5bug = 1.70506f;
constexpr float ucqcst = 0.1f, uiqcst = 0.0001f, uincst = 0.0001f;
}  // namespace mc

// Koenig (1971) ice-crystal growth parameters (setref.f90:473-600), turned
// into ckoe*10^(3 pkoe) on the host in binary32 and kept in constant memory.
__constant__ float c_ckoe[41];
__constant__ float c_pkoe[41];

// exp and log in binary32 (xmath.h: float-float internals, the correctly
// rounded result but for arguments within ~2^-40 of a rounding boundary).
// The reference's FP32 build calls glibc's expf/logf, which round within
// about 0.5 ulp; the two agree on all but ~1e-4 of arguments.
__device__ __forceinline__ float ex(float x) { return xexpf(x); }
__device__ __forceinline__ float lg(float x) { return xlogf(x); }

#define LD(arr, off) (*(const float *)((const char *)(arr) + (off)))
#define ST(arr, off, v) (*(float *)((char *)(arr) + (off)) = (v))

__device__ __forceinline__ float fminr(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmaxr(float a, float b) { return a > b ? a : b; }

// Upper bound on sedimentation sub-steps: reached only by fall speeds above
// dz*jcb*4096/dtb (kilometres per second), where the clamp keeps every lane's
// loop finite instead of following a corrupted state.
constexpr int kMaxFallSteps = 4096;
// Sub-step counts up to this run as one pipelined sweep (registers scale
// with it); larger counts fall back to one sweep per sub-step.
constexpr int kFallPipe = 8;

// Minimum wavefronts per SIMD asked of the compiler for the source kernel
// (0: its own choice, 121 VGPRs and 4 waves).  Measured at 1024x1024x80,
// ms per launch: 4 waves 4.67; 5 (96 VGPRs, 32 B/lane spilled) 4.81;
// 6 (80 VGPRs, 76 B/lane spilled) 4.76.
#ifndef CRESS_MICRO_WPE
#define CRESS_MICRO_WPE 0
#endif
#if CRESS_MICRO_WPE > 0
#define CRESS_MICRO_ATTR __attribute__((amdgpu_waves_per_eu(CRESS_MICRO_WPE)))
#else
#define CRESS_MICRO_ATTR
#endif

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) CRESS_MICRO_ATTR k_micro_src(Tiles T, Grid g, MicroK K, MicroArgs a) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 1 + xt.y * blockDim.y + threadIdx.y;
  const int k = 1 + xt.z;
  float cfl = K.dtb;
  if (i >= 1 && i <= g.ni - 1 && j <= g.nj - 1) {
    using namespace mc;
    const long p = g.at(i, j, k);
    const long qn = a.qn;
    const float thresq = K.thresq, dtb = K.dtb;
    const float rbr = MP(CRESS_RBR, a.rbr, p, k);
    // getvdens, getexner (past p')
    const float rbv = 1.0f / rbr;
    const float pr = MP(CRESS_PBR, a.pbr, p, k) + a.ppp[p];
    const float pi = ex((rd / cp) * lg((1.0f / p0) * pr));
    // past level
    const float qvp = a.qp[p], qcp = a.qpw[p], qrp = a.qpw[p + qn], qip = a.qpi[p];
    const float qsp = a.qpi[p + qn], qgp = a.qpi[p + 2 * qn], ncip = a.ncip[p];
    // diagnw / diagnsg
    const float nccp = nclcst * (1.0f / rbr);
    const float ncrp = fminr(fmaxr(sqrtf(sqrtf((nr0 * nr0 * nr0 / (cc * rhow)) * rbr * qrp)) * (1.0f / rbr),
                                   (1.e-2f / mrmax) * qrp), (1.e2f / mr0) * qrp);
    const float ncsp = fminr(fmaxr(sqrtf(sqrtf((ns0 * ns0 * ns0 / (cc * rhos)) * rbr * qsp)) * rbv,
                                   (1.e-2f / msmax) * qsp), (1.e2f / ms0) * qsp);
    const float ncgp = fminr(fmaxr(sqrtf(sqrtf((ng0 * ng0 * ng0 / (cc * rhog)) * rbr * qgp)) * rbv,
                                   (1.e-2f / mgmax) * qgp), (1.e2f / mg0) * qgp);
    // future level: the value the reference's steps (steps.f90:520-760)
    // and its fills leave at this cell.  vbcs copies levels (k=1 <- 2,
    // k=nk-1 <- nk-2).  Periodic sides: the leapfrog of the cell's periodic
    // image from bcycle (i=1 <- ni-2, i=ni-1 <- 2, the same in j).  Open
    // sides: the leapfrog inside; on the boundary ring the radiative fills
    // (rbcpt, rbcqv, rbcq, rbcs) and bc4news's corner means.
    const int ks = k == 1 ? 2 : (k == g.nk - 1 ? g.nk - 2 : k);
    float ptpf, qvf, qcf, qrf, qif, qsf, qgf, ncif;
    if (a.given) {
      ptpf = a.ptpf[p];
      qvf = a.qf[p];
      qcf = a.qfw[p];
      qrf = a.qfw[p + qn];
      qif = a.qfi[p];
      qsf = a.qfi[p + qn];
      qgf = a.qfi[p + 2 * qn];
      ncif = a.ncif[p];
    } else if (a.open && (i == 1 || i == g.ni - 1 || j == 1 || j == g.nj - 1)) {
      ptpf = obc_lf_ring(g, a.lf, false, a.ptpc, a.ptpp, i, j, ks);
      qvf = obc_lf_ring(g, a.lf, true, a.qc, a.qp, i, j, ks);
      qcf = obc_lf_ring(g, a.lf, true, a.qc + qn, a.qp + qn, i, j, ks);
      qrf = obc_lf_ring(g, a.lf, true, a.qc + 2 * qn, a.qp + 2 * qn, i, j, ks);
      qif = obc_lf_ring(g, a.lf, true, a.qc + 3 * qn, a.qp + 3 * qn, i, j, ks);
      qsf = obc_lf_ring(g, a.lf, true, a.qc + 4 * qn, a.qp + 4 * qn, i, j, ks);
      qgf = obc_lf_ring(g, a.lf, true, a.qc + 5 * qn, a.qp + 5 * qn, i, j, ks);
      ncif = obc_lf_ring(g, a.lf, false, a.ncic, a.ncip, i, j, ks);
    } else {
      const int is = a.open ? i : (i == 1 ? g.ni - 2 : (i == g.ni - 1 ? 2 : i));
      const int js = a.open ? j : (j == 1 ? g.nj - 2 : (j == g.nj - 1 ? 2 : j));
      const long ps = g.at(is, js, ks);
      const float d = dtb / MP(CRESS_RST, a.rst, ps, ks);   // dtb is the leapfrog's 2*dtb here
      ptpf = a.ptpp[ps] + a.ptfrc[ps] * d;
      qvf = fmaxr(a.qp[ps] + a.qvfrc[ps] * d, 0.0f);
      qcf = fmaxr(a.qpw[ps] + a.sfrc[ps] * d, 0.0f);
      qrf = fmaxr(a.qpw[ps + qn] + a.sfrc[ps + qn] * d, 0.0f);
      qif = fmaxr(a.qpi[ps] + a.sfrc[ps + 2 * qn] * d, 0.0f);
      qsf = fmaxr(a.qpi[ps + qn] + a.sfrc[ps + 3 * qn] * d, 0.0f);
      qgf = fmaxr(a.qpi[ps + 2 * qn] + a.sfrc[ps + 4 * qn] * d, 0.0f);
      ncif = a.ncip[ps] + a.sfrc[ps + 5 * qn] * d;
    }
    // adjstnci before the sources
    ncif = fminr(fmaxr(ncif, (0.5f / mimax) * qif), (1.e2f / mi0) * qif);
    // termblk
    const float ucq = qcp > thresq ? ucqcst : 0.0f;
    const float uiq = qip > thresq ? uiqcst : 0.0f;
    const float lnr0r = lg(r0 * rbv);
    const float urq = qrp > thresq
        ? K.curq * ex(gur * lnr0r) * ex((bur * (1.0f / 3.0f)) * lg((1.0f / (cc * rhow)) * qrp / ncrp)) : 0.0f;
    const float usq = qsp > thresq
        ? K.cusq * ex(gus * lnr0r) * ex((bus * (1.0f / 3.0f)) * lg((1.0f / (cc * rhos)) * qsp / ncsp)) : 0.0f;
    const float ugq = qgp > thresq
        ? K.cugq * ex(gug * lnr0r) * ex((bug * (1.0f / 3.0f)) * lg((1.0f / (cc * rhog)) * qgp / ncgp)) : 0.0f;
    // ---- sources: srcblk (plain), the network's tendencies, or nothing ----
    // mode 2 reads the cell's choice from k_mlp_select; every other mode
    // runs srcblk, and mode 1 also records the cell's (state, tendency)
    // pair when a process is active.
    const int how = a.mode == 2 ? a.mflag[p] : 0;
    if (how == 0) {
      const float ptpf0 = ptpf, qvf0 = qvf, qcf0 = qcf, qrf0 = qrf, qif0 = qif, qsf0 = qsf, qgf0 = qgf;
      // ---- setblk ----
      const float oned3 = 1.0f / 3.0f;
      const float t = (MP(CRESS_PTBR, a.ptbr, p, k) + a.ptpp[p]) * pi;
      const float tcel = t - t0;
      const float esw = es0 * ex(17.269f * tcel / (t - 35.86f));
      const float esi = es0 * ex(21.875f * tcel / (t - 7.66f));
      const float qvsst0 = qvp - (epsva * es0) / (pr - es0);
      const float qvsw = epsva * esw / (pr - esw);
      const float qvsi = epsva * esi / (pr - esi);
      const float lv = lv0 * ex((0.167f + 3.67e-4f * t) * lg(t0 / t));
      const float lf = lf0 + (cw - ci) * tcel;
      const float ls = lv + lf;
      const float p20dvp = p20 / pr;
      const float nu = p20dvp * ex(1.754f * lg(K.cnu * t));
      const float kp = (kp0 / (120.0f + t)) * ex(1.5f * lg((1.0f / t23) * t));
      const float mu = rbr * nu;
      const float dv = p20dvp * ex(1.81f * lg(K.cdv * t));
      const float cvnt = sqrtf(sqrtf(r0 * rbv) / nu);
      const float diaqc = qcp > thresq ? ex(oned3 * lg((6.0f / (cc * rhow)) * qcp / nccp)) : 0.0f;
      float diaqr = 0.0f, vntr = 0.0f, mi = 0.0f, diaqi = 0.0f, diaqs = 0.0f, vnts = 0.0f, diaqg = 0.0f, vntg = 0.0f;
      if (qrp > thresq) {
        diaqr = ex(oned3 * lg((1.0f / (cc * rhow)) * qrp / ncrp));
        vntr = rbr * ncrp * (0.78f * diaqr + cvnt * K.cvntr * ex(K.pdiaqr * lg(diaqr)));
      }
      if (qip > thresq) {
        mi = qip / ncip;
        diaqi = ex(oned3 * lg((6.0f / (cc * rhoi)) * mi));
      }
      if (qsp > thresq) {
        diaqs = ex(oned3 * lg((1.0f / (cc * rhos)) * qsp / ncsp));
        vnts = rbr * ncsp * (0.78f * diaqs + cvnt * K.cvnts * ex(K.pdiaqs * lg(diaqs)));
      }
      if (qgp > thresq) {
        diaqg = ex(oned3 * lg((1.0f / (cc * rhog)) * qgp / ncgp));
        vntg = rbr * ncgp * (0.78f * diaqg + cvnt * K.cvntg * ex(K.pdiaqg * lg(diaqg)));
      }
      // ---- nuc1stv ----
      float nuvi = 0.0f;
      if (qvp > thresq && qvp > qvsi && t > tlow && t < t0) {
        const float qvssi = qvp - qvsi;
        const float aa = 15.25f * qvp / qvsi - 10.08f;
        nuvi = aa < 40.0f ? fmaxr(0.0f, fminr(mi0 * ex(aa) * rbv - qip, qvssi)) : qvssi;
      }
      // ---- nuc1stc ----
      float nuci = 0.0f;
      if (qcp > thresq) {
        const float nufci = (t > tlow && t < t0)
            ? fminr(K.rwdt2 * (ex(-0.66f * tcel) - 1.0f) * qcp * qcp / nccp, qcp) : 0.0f;
        float nucci = 0.0f;
        if (t > tlow && t < 270.16f) {
          const float piv = 1.0f / pr;
          const float knd = K.cknd * piv * t;
          const float dar = K.cdar * t * (1.0f + knd) / mu;
          const float f1 = K.cc45 * ex(1.3f * lg(270.16f - t)) * diaqc;
          // the reference's contact branch compares tc with t where tc=t, so
          // its thermophoretic term is always zero (nuc1stc.f90:334-360)
          nucci = (0.0f + f1 * dar) * qcp * dtb;
        }
        const float nuhci = t <= tlow ? qcp : 0.0f;
        nuci = nufci + nucci + nuhci;
      }
      // ---- collect ----
      const float r0rsq = sqrtf(r0 * rbr);
      float diaqr2 = 0.0f, diaqr3 = 0.0f, qr2b = 0.0f, diaqs2 = 0.0f, diaqs3 = 0.0f, qs2b = 0.0f;
      float diaqg2 = 0.0f, qg2b = 0.0f;
      if (qrp > thresq) { diaqr2 = diaqr * diaqr; diaqr3 = diaqr * diaqr2; qr2b = r0rsq * ex((2.0f + bur) * lg(diaqr)); }
      if (qsp > thresq) { diaqs2 = diaqs * diaqs; diaqs3 = diaqs * diaqs2; qs2b = r0rsq * ex((2.0f + bus) * lg(diaqs)); }
      if (qgp > thresq) { diaqg2 = diaqg * diaqg; qg2b = r0rsq * ex((2.0f + bug) * lg(diaqg)); }
      float clcr = 0.0f, clcs = 0.0f, clcg = 0.0f, clri = 0.0f, clrs = 0.0f, clrg = 0.0f, clir = 0.0f, clsr = 0.0f;
      float clis = 0.0f, clig = 0.0f, clsg = 0.0f, aa, bb, sink;
      if (t > tlow) {
        if (qcp > thresq) {
          const float cstk = ((1.0f / 9.0f) * rhow) * diaqc * diaqc;
          if (qrp > thresq) {
            aa = cstk * urq / (diaqr * mu);
            bb = aa + 0.5f;
            aa = aa * aa / (bb * bb);
            clcr = K.cclcr * aa * qr2b * ncrp * qcp;
          }
          if (qsp > thresq) {
            aa = cstk * usq / (diaqs * mu);
            bb = aa + 0.5f;
            const float ecs = aa * aa / (bb * bb);
            clcs = K.cclcs * qs2b * ecs * ncsp * qcp;
          }
          if (qgp > thresq) {
            aa = cstk * ugq / (diaqg * mu);
            bb = aa + 0.5f;
            aa = aa * aa / (bb * bb);
            clcg = K.cclcg * aa * qg2b * ncgp * qcp;
          }
          sink = clcr + clcs + clcg;
          if (qcp < sink) { aa = qcp / sink; clcr = clcr * aa; clcs = clcs * aa; clcg = clcg * aa; }
        }
        if (qrp > thresq) {
          if (qip > thresq && t < t0) { clri = K.cclri * qr2b * ncrp * qip; clir = clri; }
          if (qsp > thresq) {
            aa = 2.0f * diaqr * diaqs;
            bb = urq - usq;
            bb = ncrp * ncsp * rbr * sqrtf(bb * bb + 0.04f * urq * usq);
            clrs = bb * K.cclrs * (5.0f * diaqr2 + aa + 0.5f * diaqs2) * diaqr3;
            clsr = bb * K.cclsr * (5.0f * diaqs2 + aa + 0.5f * diaqr2) * diaqs3;
          }
          if (qgp > thresq) {
            aa = 5.0f * diaqr2 + 2.0f * diaqr * diaqg + 0.5f * diaqg2;
            bb = urq - ugq;
            clrg = aa * K.cclrg * diaqr3 * ncrp * ncgp * rbr * sqrtf(bb * bb + 0.04f * urq * ugq);
          }
          sink = clri + clrs + clrg;
          if (qrp < sink) { aa = qrp / sink; clri = clri * aa; clrs = clrs * aa; clrg = clrg * aa; }
        }
      }
      if (qip > thresq && t < t0) {
        if (qsp > thresq) clis = K.cclis * qs2b * ncsp * qip;
        if (qgp > thresq) clig = K.cclig * qg2b * ncgp * qip;
        sink = clir + clis + clig;
        if (qip < sink) { aa = qip / sink; clir = clir * aa; clis = clis * aa; clig = clig * aa; }
      }
      if (qsp > thresq) {
        if (qgp > thresq) {
          aa = 5.0f * diaqs2 + 2.0f * diaqs * diaqg + 0.5f * diaqg2;
          bb = usq - ugq;
          clsg = aa * (t < t0 ? K.cclsg : K.cclsg3) * diaqs3 * ncsp * ncgp * rbr * sqrtf(bb * bb + 0.04f * usq * ugq);
        }
        sink = clsr + clsg;
        if (qsp < sink) { aa = qsp / sink; clsr = clsr * aa; clsg = clsg * aa; }
      }
      // ---- prodctwg ----
      float pgwet = -1.0f;
      if (qgp > thresq && tcel < t0cel) {
        const float pgdry = clcg + clrg + clig + clsg;
        const float lfice = lf + cw * tcel;
        const float cligw = (1.0f / eig) * clig, clsgw = (1.0f / esg) * clsg;
        const float pg = K.cc2dtn * vntg * (lv * dv * rbr * qvsst0 + kp * tcel) / (lfice * rbr) +
                         (cligw + clsgw) * (1.0f - ci * tcel / lfice);
        if (pg > 0.0f && pg < pgdry) {
          pgwet = pg;
          clig = cligw;
          clsg = clsgw;
          sink = clir + clis + clig;
          if (qip < sink) { aa = qip / sink; clir = clir * aa; clis = clis * aa; clig = clig * aa; }
          sink = clsr + clsg;
          if (qsp < sink) { aa = qsp / sink; clsr = clsr * aa; clsg = clsg * aa; }
        }
      }
      // ---- distrpg ----
      float clrsg = 0.0f;
      if (qsp > thresq) {
        if (t < t0) {
          aa = diaqs * diaqs;
          bb = diaqr * diaqr;
          aa = (rhos * rhos) * aa * aa * aa;
          bb = (rhow * rhow) * bb * bb * bb;
          const float alpha = aa / (aa + bb), alpha1 = 1.0f - alpha;
          clrsg = alpha1 * clrs;
          clrs = alpha * clrs;
          clsr = alpha1 * clsr;
        } else {
          clrsg = clrs;
          clrs = 0.0f;
        }
      }
      // ---- aggregat ----
      const float agin = qip > thresq ? rbr * ncip * qip * ex(oned3 * lg(K.cagin_ag * rbv)) : 0.0f;
      // ---- melting ----
      const float cmlxr1 = K.cc2dt * rbv * (kp * tcel + lv * dv * rbr * qvsst0);
      const float cmlxr2 = cw * tcel;
      float mlic = (qip > thresq && tcel >= t0cel) ? qip : 0.0f;
      float mlsr = (qsp > thresq && tcel >= t0cel) ? fmaxr((cmlxr1 * vnts + cmlxr2 * (clcs + clrs)) / lf, 0.0f) : 0.0f;
      float mlgr = (qgp > thresq && tcel >= t0cel) ? fmaxr((cmlxr1 * vntg + cmlxr2 * (clcg + clrg)) / lf, 0.0f) : 0.0f;
      // ---- depsit ----
      float vdvr = 0.0f, vdvi = 0.0f, vdvs = 0.0f, vdvg = 0.0f;
      if (t > tlow) {
        const float qvssw = qvp - qvsw, qvssi = qvp - qvsi;
        const float ad = 1.0f / (rv * kp * t * t);
        const float bd = dv * rbr;
        const float cd = K.ccdtb4 * rbv;
        const float gi = 1.0f / (ls * ls * ad + 1.0f / (bd * qvsi));
        const float cvdvx1 = cd * (qvp / qvsw - 1.0f) / (lv * lv * ad + 1.0f / (bd * qvsw));
        const float cvdvx2 = K.ccdtb4 * dv * qvsst0;
        const float cvdvx3 = cd * (qvp / qvsi - 1.0f);
        const float cvdvx4 = ad * ls * lf;
        if (qrp > thresq && qvp < qvsw) vdvr = fmaxr(cvdvx1 * vntr, qvssw);
        if (qip > thresq && t < t0 - 0.01f) {
          const int itsc = (int)(-tcel);
          vdvi = c_ckoe[itsc] * (qvp - qvsi) / (qvsw - qvsi) * ncip * ex(c_pkoe[itsc] * lg(mi)) * dtb;
        }
        if (qsp > thresq) {
          if (t > t0) vdvs = mlsr > 0.0f ? cvdvx2 * vnts : cvdvx1 * vnts;
          else vdvs = gi * (cvdvx3 * vnts - cvdvx4 * clcs);
        }
        if (qgp > thresq) {
          if (t > t0) vdvg = mlgr > 0.0f ? cvdvx2 * vntg : cvdvx1 * vntg;
          else vdvg = gi * (cvdvx3 * vntg - cvdvx4 * clcg);
        }
        sink = vdvi + vdvs + vdvg;
        if ((qvssi < sink && qvssi > 0.0f) || (qvssi > sink && qvssi < 0.0f)) {
          const float h = qvssi / sink;
          vdvi = vdvi * h; vdvs = vdvs * h; vdvg = vdvg * h;
        }
      }
      // ---- convers ----
      float cncr = 0.0f, cnis = 0.0f, cnsg = 0.0f;
      if (qcp > thresq && t > tlow && qcp >= K.qccm * nccp)
        cncr = K.ccncr * rbr * ex(-oned3 * lg(nccp)) * ex((7.0f / 3.0f) * lg(qcp)) / mu;
      if (t < t0) {
        if (qip > thresq) {
          float cn = rbr * ex(oned3 * lg(K.cagin_cv * rbv)) * qip * qip / lg(K.diaqs0 / diaqi);
          if (vdvi >= 0.0f) {
            const float ms05 = 0.5f * ms0;
            if (mi < ms05) cn = cn + mi / (ms0 - mi) * vdvi;
            else cn = cn + (vdvi + (1.0f - ms05 / mi) * qip);
          }
          cnis = fminr(cn, qip);
        }
        if (qsp > thresq && vdvs > 0.0f && vdvs < clcs) cnsg = fminr(K.ccnsg * clcs, qsp);
      }
      // ---- nuc2nd ----
      float spsi = 0.0f, spgi = 0.0f;
      if (t <= 270.16f && t > 265.16f) {
        const float fr = t > 268.16f ? (270.16f - t) * K.mi0352 * rbv : (t - 265.16f) * K.mi0353 * rbv;
        spsi = fr * clcs;
        spgi = pgwet > 0.0f ? 0.0f : fr * clcg;
      }
      // ---- freezing ----
      float frrg = 0.0f;
      if (qrp > thresq && tcel < t0cel) {
        if (tcel <= tlow - t0) {
          frrg = qrp;
        } else {
          const float d3 = diaqr * diaqr * diaqr;
          frrg = fminr(qrp, K.cfrrg * d3 * d3 * (ex(-0.66f * tcel) - 1.0f) * ncrp);
        }
      }
      // ---- shedding ----
      float shsr = (qsp > thresq && t >= t0) ? clcs + clrs : 0.0f;
      float shgr = 0.0f;
      if (qgp > thresq) {
        if (t >= t0) shgr = clcg + clrg;
        else if (pgwet > 0.0f) shgr = (clcg + clrg + clig + clsg) - pgwet;
      }
      // ---- more0q ----
      float h;
      if (qcp > thresq) {
        sink = nuci + clcr + clcs + clcg + cncr - mlic;
        if (qcf < sink) {
          h = qcf / sink;
          nuci *= h; clcr *= h; clcs *= h; clcg *= h; cncr *= h; mlic *= h;
        }
      }
      if (qip > thresq) {
        sink = clir + clis + clig + cnis + mlic - vdvi - nuvi - nuci - spsi - spgi;
        if (qif < sink) {
          h = qif / sink;
          clir *= h; clis *= h; clig *= h; cnis *= h; mlic *= h; vdvi *= h; nuvi *= h; nuci *= h; spsi *= h; spgi *= h;
        }
      }
      if (qrp > thresq) {
        sink = clri + clrs + clrg + clrsg + frrg - vdvr - clcr - cncr - mlsr - mlgr - shsr - shgr;
        if (qrf < sink) {
          h = qrf / sink;
          clri *= h; clrs *= h; clrg *= h; clrsg *= h; frrg *= h; vdvr *= h; clcr *= h; cncr *= h;
          mlsr *= h; mlgr *= h; shsr *= h; shgr *= h;
        }
      }
      if (qsp > thresq) {
        sink = clsr + clsg + cnsg + spsi + mlsr + shsr - vdvs - clcs - clrs - clis - cnis;
        if (qsf < sink) {
          h = qsf / sink;
          clsr *= h; clsg *= h; cnsg *= h; spsi *= h; mlsr *= h; shsr *= h; vdvs *= h; clcs *= h; clrs *= h;
          clis *= h; cnis *= h;
        }
      }
      if (qgp > thresq) {
        sink = spgi + mlgr + shgr - vdvg - clri - clir - clsr - clcg - clrg - clig - clsg - clrsg - cnsg - frrg;
        if (qgf < sink) {
          h = qgf / sink;
          spgi *= h; mlgr *= h; shgr *= h; vdvg *= h; clri *= h; clir *= h; clsr *= h; clcg *= h; clrg *= h;
          clig *= h; clsg *= h; clrsg *= h; cnsg *= h; frrg *= h;
        }
      }
      // ---- newblk ----
      const float cppiv = 1.0f / (cp * pi);
      const float clix = clir + clis + clig;
      const float spxi = spsi + spgi;
      const float cxcr = clcr + cncr;
      const float cxsg = clsr + clsg + cnsg;
      const float nmci = nuci - mlic;
      const float mssr = mlsr + shsr;
      const float msgr = mlgr + shgr;
      const float nuvdvx = vdvi + vdvs + vdvg + nuvi;
      const float clnmci = clcs + clcg + nmci;
      const float cfmsxr = clri + clrs + clrg + clrsg + frrg - mssr - msgr;
      ptpf = ptpf + (vdvr * lv + (clnmci + cfmsxr) * lf + nuvdvx * ls) * cppiv;
      float qvsink = nuvdvx + vdvr;
      float qxsink = cfmsxr - cxcr - vdvr;
      if (qrf < qxsink) {
        qxsink = cxcr + clnmci + (qxsink - qrf);
        qrf = 0.0f;
      } else {
        qrf = qrf - qxsink;
        qxsink = cxcr + clnmci;
      }
      if (qcf < qxsink) {
        const float dqv = qxsink - qcf;
        ptpf = ptpf + cppiv * dqv * lv;
        qvsink = qvsink + dqv;
        qcf = 0.0f;
      } else {
        qcf = qcf - qxsink;
      }
      qxsink = clix - spxi - nmci + cnis - vdvi - nuvi;
      if (qif < qxsink) {
        const float dqv = qxsink - qif;
        ptpf = ptpf + cppiv * dqv * ls;
        qvsink = qvsink + dqv;
        qif = 0.0f;
      } else {
        qif = qif - qxsink;
      }
      qvf = fmaxr(qvf - qvsink, 0.0f);
      qsf = fmaxr(qsf - (cxsg + mssr + spsi - vdvs - clcs - clrs - clis - cnis), 0.0f);
      qgf = fmaxr(qgf + (cxsg - msgr - spgi + vdvg + clri + clir + clcg + clrg + clig + clrsg + frrg), 0.0f);
      if (qcp > thresq) ncif = ncif + nuci * nccp / qcp;
      if (qip > thresq) {
        if (vdvi < 0.0f)
          ncif = ncif - (agin + (clix + mlic - vdvi) / mi - (spxi + nuvi) * (1.0f / mi0) + cnis * (1.0f / ms0));
        else
          ncif = ncif - (agin + (clix + mlic) / mi - (spxi + nuvi) * (1.0f / mi0) + cnis * (1.0f / ms0));
      } else {
        ncif = ncif + (spxi + nuvi) * (1.0f / mi0);
      }
      if (a.mode == 1) {
        const bool act = qcp > thresq || qrp > thresq || qip > thresq || qsp > thresq || qgp > thresq || nuvi > 0.0f;
        const unsigned long long ball = __ballot(act);
        if (act) {
          const int lane = (int)(threadIdx.x + threadIdx.y * blockDim.x) & 63;
          const int lead = __ffsll((long long)ball) - 1;
          unsigned long long base = 0;
          if (lane == lead) base = atomicAdd(a.cap.count, (unsigned long long)__popcll(ball));
          base = __shfl(base, lead);
          const unsigned long long idx = base + __popcll(ball & ((1ull << lane) - 1ull));
          if (idx < (unsigned long long)a.cap.cap) {
            float *fi = a.cap.in + idx * MLP_NIN, *ft = a.cap.tgt + idx * MLP_NOUT;
            fi[0] = t; fi[1] = pr; fi[2] = rbr;
            fi[3] = qvp; fi[4] = qcp; fi[5] = qrp; fi[6] = qip; fi[7] = qsp; fi[8] = qgp;
            const float rdt = 1.0f / dtb;
            ft[0] = (ptpf - ptpf0) * pi * rdt;
            ft[1] = (qvf - qvf0) * rdt; ft[2] = (qcf - qcf0) * rdt; ft[3] = (qrf - qrf0) * rdt;
            ft[4] = (qif - qif0) * rdt; ft[5] = (qsf - qsf0) * rdt; ft[6] = (qgf - qgf0) * rdt;
          }
        }
      }
    } else if (how == 1) {
      // the network's tendencies over the step 2 dtb, then the hard
      // constraints of newblk.f90:416-489: no species below zero, total
      // water conserved (vapour takes the balance of the hydrometeor
      // changes, which are scaled down when the vapour cannot), and the
      // latent heating that the phase changes imply,
      //   cp pi dtheta = Lv (-dqv) + Lf (dqi + dqs + dqg),
      // with Lv, Lf at the cell's temperature as setblk forms them.
      const float t = (MP(CRESS_PTBR, a.ptbr, p, k) + a.ptpp[p]) * pi;
      const float lv = lv0 * ex((0.167f + 3.67e-4f * t) * lg(t0 / t));
      const float lf = lf0 + (cw - ci) * (t - t0);
      const long qn7 = a.qn;
      float dqc = fmaxr(a.mout[2 * qn7 + p] * dtb, -qcf);
      float dqr = fmaxr(a.mout[3 * qn7 + p] * dtb, -qrf);
      float dqi = fmaxr(a.mout[4 * qn7 + p] * dtb, -qif);
      float dqs = fmaxr(a.mout[5 * qn7 + p] * dtb, -qsf);
      float dqg = fmaxr(a.mout[6 * qn7 + p] * dtb, -qgf);
      float dqh = dqc + dqr + dqi + dqs + dqg;
      if (dqh > qvf) {
        const float sc = qvf / dqh;
        dqc *= sc; dqr *= sc; dqi *= sc; dqs *= sc; dqg *= sc;
        dqh = dqc + dqr + dqi + dqs + dqg;
      }
      const float dqv = -dqh;
      ptpf = ptpf + (lv * dqh + lf * (dqi + dqs + dqg)) * (1.0f / (cp * pi));
      qvf = fmaxr(qvf + dqv, 0.0f);
      qcf = fmaxr(qcf + dqc, 0.0f);
      qrf = fmaxr(qrf + dqr, 0.0f);
      qif = fmaxr(qif + dqi, 0.0f);
      qsf = fmaxr(qsf + dqs, 0.0f);
      qgf = fmaxr(qgf + dqg, 0.0f);
    }
    // adjstnci after the sources
    ncif = fminr(fmaxr(ncif, (0.5f / mimax) * qif), (1.e2f / mi0) * qif);
    a.ptpf[p] = ptpf;
    a.qf[p] = qvf;
    a.qfw[p] = qcf;
    a.qfw[p + qn] = qrf;
    a.qfi[p] = qif;
    a.qfi[p + qn] = qsf;
    a.qfi[p + 2 * qn] = qgf;
    a.ncif[p] = ncif;
    a.urq[p] = urq;
    a.usq[p] = usq;
    a.ugq[p] = ugq;
    // sedimentation CFL (fallblk.f90:364-377)
    const float dzjcb = K.dz * MP(CRESS_JCB, a.jcb, p, k);
    cfl = fminr(dzjcb / (ucq + eps), cfl);
    cfl = fminr(dzjcb / (urq + eps), cfl);
    cfl = fminr(dzjcb / (uiq + eps), cfl);
    cfl = fminr(dzjcb / (usq + eps), cfl);
    cfl = fminr(dzjcb / (ugq + eps), cfl);
  }
  // block minimum, then one atomic per wavefront-reduced block
  __shared__ float red[256];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  red[tid] = cfl;
  __syncthreads();
  for (int s = (blockDim.x * blockDim.y) / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fminr(red[tid], red[tid + s]);
    __syncthreads();
  }
  if (tid == 0) atomicMin(a.dtpmin, __float_as_uint(red[0]));
}

// ---------------------------------------------------------------------------
// The network's cells for this step (MlpSelectArgs in micro.h).  Same cells
// as k_micro_src.  The activity test is srcblk's own gating: every process
// rate is zero unless a hydrometeor exceeds thresq or nuc1stv fires
// (qv > thresq, qv > qvsi, tlow < T < t0; nuc1stv.f90, setblk.f90:333-339).
__global__ void __launch_bounds__(256) k_mlp_select(Tiles T, Grid g, MlpSelectArgs a) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 1 + xt.y * blockDim.y + threadIdx.y;
  const int k = 1 + xt.z;
  bool net = false;
  float f[MLP_NIN];
  long p = 0;
  if (i >= 1 && i <= g.ni - 1 && j <= g.nj - 1) {
    using namespace mc;
    p = g.at(i, j, k);
    const long qn = a.qn;
    const float thresq = a.thresq;
    const float rbr = MP(CRESS_RBR, a.rbr, p, k);
    const float pr = MP(CRESS_PBR, a.pbr, p, k) + a.ppp[p];
    const float pi = ex((rd / cp) * lg((1.0f / p0) * pr));
    const float t = (MP(CRESS_PTBR, a.ptbr, p, k) + a.ptpp[p]) * pi;
    const float qvp = a.qp[p], qcp = a.qpw[p], qrp = a.qpw[p + qn], qip = a.qpi[p];
    const float qsp = a.qpi[p + qn], qgp = a.qpi[p + 2 * qn];
    const float esi = es0 * ex(21.875f * (t - t0) / (t - 7.66f));
    const float qvsi = epsva * esi / (pr - esi);
    const bool act = qcp > thresq || qrp > thresq || qip > thresq || qsp > thresq || qgp > thresq ||
                     (qvp > thresq && qvp > qvsi && t > tlow && t < t0);
    f[0] = t; f[1] = pr; f[2] = rbr; f[3] = qvp; f[4] = qcp; f[5] = qrp; f[6] = qip; f[7] = qsp; f[8] = qgp;
    bool inside = true;
    for (int c = 0; c < MLP_NIN; c++) inside = inside && f[c] >= a.norm[64 + c] && f[c] <= a.norm[80 + c];
    net = act && inside;
    a.s.flag[p] = !act ? 2 : (inside ? 1 : 0);
  }
  const unsigned long long ball = __ballot(net);
  if (net) {
    const int lane = (int)(threadIdx.x + threadIdx.y * blockDim.x) & 63;
    const int lead = __ffsll((long long)ball) - 1;
    unsigned base = 0;
    if (lane == lead) base = atomicAdd(a.s.count, (unsigned)__popcll(ball));
    base = __shfl(base, lead);
    const unsigned idx = base + (unsigned)__popcll(ball & ((1ull << lane) - 1ull));
    a.s.list[idx] = (int)p;
    unsigned short *x = a.s.x + (long)idx * MLP_INP;
    for (int c = 0; c < MLP_INP; c++)
      x[c] = c < MLP_NIN ? f2bf((f[c] - a.norm[c]) * a.norm[16 + c]) : (unsigned short)0;
  }
}

// fallblk.f90:439-541 on the columns (i,j), i in 1..ni-1, j = 1+tile y,
// one wavefront per species (threadIdx.y: qc, qi, nci, qr, qs, qg), each
// doing its npstp upwind sub-steps (upwqp.f90:164-202 with the surface
// precipitation rate and accumulation; upwnp.f90:143-176 for nci); then the
// trailing adjstnci (coldblk.f90:463-465) once qi is final.  Species are
// independent in fallblk, so the reference's species order does not matter.
__global__ void __launch_bounds__(384) k_micro_fall(Tiles T, Grid g, FallArgs a) {
  const uint3 xt = xcd_tile(T);
  const int i = CRESS_FALL_ORG + xt.x * 64 + threadIdx.x;
  const int j = 1 + xt.y;
  const int sp = threadIdx.y;
  const bool in = i >= 1 && i <= g.ni - 1;
  const float dtb = a.dtb;
  float dtp = fminr(__uint_as_float(*a.dtpmin), dtb);
  int npstp;
  if (dtp < dtb) {
    const float r = (dtb + 0.001f) / dtp;
    npstp = r < (float)(kMaxFallSteps - 1) ? (int)r + 1 : kMaxFallSteps;
    dtp = dtb / (float)npstp;
  } else {
    npstp = 1;
    dtp = dtb;
  }
  const uint32_t sk = 4u * (uint32_t)g.nxy;
  const uint32_t q1 = 4u * (uint32_t)g.at(i, j, 1);
  const long qn = a.qn;
  if (in) {
    // species table: the future array, its fall speed and precipitation;
    // cloud water and ice fall at constant speeds where present on the past
    // level (termblk.f90:303-330)
    float *q;
    const float *ufield = nullptr, *qpres = nullptr;
    float ucst = 0.0f, *prec = nullptr;
    switch (sp) {
      case 0: q = a.qfw; qpres = a.qpw; ucst = mc::ucqcst; prec = a.prwtr; break;
      case 1: q = a.qfi; qpres = a.qpi; ucst = mc::uiqcst; prec = a.price; break;
      case 2: q = a.ncif; qpres = a.qpi; ucst = mc::uincst; break;
      case 3: q = a.qfw + qn; ufield = a.urq; prec = a.prwtr + 2 * a.n2; break;
      case 4: q = a.qfi + qn; ufield = a.usq; prec = a.price + 2 * a.n2; break;
      default: q = a.qfi + 2 * qn; ufield = a.ugq; prec = a.price + 4 * a.n2; break;
    }
    const float thresq = a.thresq;
    auto speed = [&](uint32_t o) {
      return ufield ? LD(ufield, o) : (LD(qpres, o) > thresq ? ucst : 0.0f);
    };
    const float dzvdt = a.dziv * dtp;
    const int nk = g.nk;
    if (npstp <= kFallPipe) {
      // All npstp sub-steps in one upward sweep, skewed by one level per
      // sub-step: sub-step s updates level m-s+1 at iteration m, from
      // sub-step s-1's value one level up, which the same iteration has
      // just produced.  Only the last sub-step's values reach memory.
      // lo_q/lo_f: q and flux of sub-step s's input at its current level;
      // out: sub-step s's newest output (kept when s passes level nk-2,
      // which makes it the top copy q(nk-1) = q(nk-2)); cw/rw: c = rbr*u
      // one level above, and rst at, sub-step s's current level.
      float lo_q[kFallPipe], lo_f[kFallPipe], out[kFallPipe], cw[kFallPipe], rw[kFallPipe];
      const float c1 = MQ(CRESS_RBR, a.rbr, q1, 1) * speed(q1);
      const float q01 = LD(q, q1);
      lo_q[0] = q01;
      lo_f[0] = c1 * q01;
#pragma unroll
      for (int s = 0; s < kFallPipe; s++) {
        if (s > 0) {
          lo_q[s] = 0.0f;
          lo_f[s] = 0.0f;
        }
        out[s] = 0.0f;
        cw[s] = 0.0f;
        rw[s] = 0.0f;
      }
      const long c2 = (long)i + (long)(g.ni + 2) * j;   // 2-D precipitation arrays
      float acc = prec ? prec[c2 + a.n2] : 0.0f;
      float rate = 0.0f;
      const int mlast = nk - 2 + npstp - 1;
      // the inputs of level m are loaded one iteration ahead (rbr, fall
      // speed and q at m+1, rst at m), so each lane keeps two levels of
      // loads in flight; writes trail at level <= m, never ahead of them
      float nrb = 0.0f, nsp = 0.0f, nrs = 0.0f, nq = 0.0f;
      auto fetch = [&](int m) {
        const uint32_t om = q1 + (uint32_t)(m - 1) * sk;
        nrb = MQ(CRESS_RBR, a.rbr, om + sk, m + 1);
        nsp = speed(om + sk);
        nrs = MQ(CRESS_RST, a.rst, om, m);
        nq = LD(q, om + sk);
      };
      fetch(1);
      for (int m = 1; m <= mlast; m++) {
        // shift the coefficient windows one level and take level m+1 / m
#pragma unroll
        for (int s = kFallPipe - 1; s > 0; s--) {
          cw[s] = cw[s - 1];
          rw[s] = rw[s - 1];
        }
        float hi = 0.0f;
        if (m <= nk - 2) {
          cw[0] = nrb * nsp;
          rw[0] = nrs;
          hi = nq;   // q_0(m+1)
          if (m + 1 <= nk - 2) fetch(m + 1);
        }
#pragma unroll
        for (int s = 0; s < kFallPipe; s++) {
          if (s >= npstp) break;
          const int k = m - s;   // level of sub-step s+1 at this iteration
          if (k >= 1 && k <= nk - 2) {
            const float qin = s == 0 ? hi : out[s - 1];
            const float fhi = cw[s] * qin;
            const float qn = fmaxr(lo_q[s] + (fhi - lo_f[s]) / rw[s] * dzvdt, 0.0f);
            if (k == 1) {
              // first update of this sub-step: its surface rate
              rate = (lo_f[s] + fhi) * (0.5f / mc::rhow);
              acc = acc + rate * (0.5f * dtp);
            }
            lo_q[s] = qin;
            lo_f[s] = fhi;
            out[s] = qn;
            if (s + 1 < kFallPipe && k == 1) {
              lo_q[s + 1] = qn;
              lo_f[s + 1] = c1 * qn;
            }
            if (s == npstp - 1) ST(q, q1 + (uint32_t)(k - 1) * sk, qn);
          }
        }
      }
      ST(q, q1 + (uint32_t)(nk - 2) * sk, out[npstp - 1]);
      if (prec) {
        prec[c2] = rate;
        prec[c2 + a.n2] = acc;
      }
    } else {
      for (int s = 0; s < npstp; s++) {
        uint32_t o = q1;
        float q_k = LD(q, o);
        float flx_k = MQ(CRESS_RBR, a.rbr, o, 1) * speed(o) * q_k;
        const float flx1 = flx_k;
        float flx2 = 0.0f;
        for (int k = 1; k <= nk - 2; k++, o += sk) {
          const float q_kp1 = LD(q, o + sk);
          const float flx_kp1 = MQ(CRESS_RBR, a.rbr, o + sk, k + 1) * speed(o + sk) * q_kp1;
          if (k == 1) flx2 = flx_kp1;
          q_k = fmaxr(q_k + (flx_kp1 - flx_k) / MQ(CRESS_RST, a.rst, o, k) * dzvdt, 0.0f);
          ST(q, o, q_k);
          flx_k = flx_kp1;
          q_k = q_kp1;
        }
        // top copy: q(nk-1) = q(nk-2) after the update
        ST(q, q1 + (uint32_t)(nk - 2) * sk, LD(q, q1 + (uint32_t)(nk - 3) * sk));
        if (prec) {
          const long c2 = (long)i + (long)(g.ni + 2) * j;
          const float rate = (flx1 + flx2) * (0.5f / mc::rhow);
          prec[c2] = rate;
          prec[c2 + a.n2] = prec[c2 + a.n2] + rate * (0.5f * dtp);
        }
      }
    }
  }
  // adjstnci is pointwise: the six wavefronts share the column's levels
  __syncthreads();
  if (in) {
    const float *qi = a.qfi;
    for (int k = 1 + sp; k <= g.nk - 1; k += (int)blockDim.y) {
      const uint32_t o = q1 + (uint32_t)(k - 1) * sk;
      const float qv = LD(qi, o);
      ST(a.ncif, o, fminr(fmaxr(LD(a.ncif, o), (0.5f / mc::mimax) * qv), (1.e2f / mc::mi0) * qv));
    }
  }
}

// s_satadjst for one cell: getexner (getexner.f90:123-131), siadjst
// (siadjst.f90:159-269) and swadjst (swadjst.f90:237-310, cphopt<=3), on
// the cell's theta', qv, qc, qi, nci in registers.
__device__ __forceinline__ void satadj_cell(float thresq, float pbr, float pp, float tb, float &th, float &qv,
                                            float &qc, float &qi, float &ni) {
  using namespace mc;
  const float pr = pbr + pp;
  const float pi = ex((rd / cp) * lg((1.0f / p0) * pr));
  // ice
  float t = (tb + th) * pi;
  if (t <= tlow) {
    for (int pass = 0; pass < 2; pass++) {
      if (pass == 1) {
        t = (tb + th) * pi;
        if (!(t <= tlow)) break;
      }
      const float tcel = t - t0;
      const float aa = 1.0f / (t - 7.66f);
      const float bb = aa * tcel;
      const float esi = es0 * ex(21.875f * bb);
      const float qvsi = epsva * esi / (pr - esi);
      if (qi > thresq || qv > qvsi) {
        const float lscpi = (lv0 * ex((0.167f + 3.67e-4f * t) * lg(t0 / t)) + (lf0 + (cw - ci) * tcel)) / (cp * pi);
        const float dqi = (qvsi - qv) / (1.0f + 21.875f * aa * (1.0f - bb) * qvsi * lscpi * pi);
        if (qi > dqi) {
          ni = qi > thresq ? ni - dqi * ni / qi : ni - dqi * (1.0f / mi0);
          th = th - dqi * lscpi;
          qv = qv + dqi;
          qi = qi - dqi;
        } else {
          ni = 0.0f;
          th = th - qi * lscpi;
          qv = qv + qi;
          qi = 0.0f;
        }
      }
    }
  }
  // water
  for (int pass = 0; pass < 2; pass++) {
    t = (tb + th) * pi;
    const float aa = 1.0f / (t - 35.86f);
    const float bb = aa * (t - t0);
    const float esw = es0 * ex(17.269f * bb);
    const float qvsw = epsva * esw / (pr - esw);
    if (!(qc > thresq || qv > qvsw)) break;
    const float lvcpi = lv0 * ex((0.167f + 3.67e-4f * t) * lg(t0 / t)) / (cp * pi);
    const float dqc = (qvsw - qv) / (1.0f + 17.269f * aa * (1.0f - bb) * qvsw * lvcpi * pi);
    if (qc > dqc) {
      th = th - dqc * lvcpi;
      qv = qv + dqc;
      qc = qc - dqc;
    } else {
      th = th - qc * lvcpi;
      qv = qv + qc;
      qc = 0.0f;
    }
  }
}

// s_satadjst on one level over i=1..ni-1, j=1..nj-1, k=1..nk-1.
__global__ void k_satadjst(Tiles T, Grid g, float thresq, const float *pbr, const float *ptbr, const float *pp,
                           float *ptp, float *q, float *nci, long qn) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 1 + xt.y * blockDim.y + threadIdx.y;
  const int k = 1 + xt.z;
  if (i < 1 || i > g.ni - 1 || j > g.nj - 1) return;
  const long p = g.at(i, j, k);
  float th = ptp[p], qv = q[p], qc = q[p + qn], qi = q[p + 3 * qn], ni = nci[p];
  satadj_cell(thresq, MP(CRESS_PBR, pbr, p, k), pp[p], MP(CRESS_PTBR, ptbr, p, k), th, qv, qc, qi, ni);
  ptp[p] = th;
  q[p] = qv;
  q[p + qn] = qc;
  q[p + 3 * qn] = qi;
  nci[p] = ni;
}

// s_timeflt (timeflt.f90:348-484) for the twelve prognostics of one cell,
// x = (1-2c) x + c (x_f + x_p) over each field's range (u to i=ni, v to
// j=nj, w to k=nk, the rest to ni-1, nj-1, nk-1), followed, inside
// i<=ni-1, j<=nj-1, k<=nk-1, by s_satadjst on the filtered present level
// and on the future level (intgdrv.f90:968-998).  Every step is pointwise,
// so doing them per cell in this order is the reference's result.
// Fields: u, v, w, pp, ptp, qv, qc, qr, qi, qs, qg, nci.
__global__ void __launch_bounds__(256) k_asselin_sat(Tiles T, Grid g, float m1fc2, float filcoe, float thresq,
                                                     AsselinSat a) {
  const uint3 xt = xcd_tile(T);
  const int i = xt.x * blockDim.x + threadIdx.x;
  const int j = 1 + xt.y * blockDim.y + threadIdx.y;
  const int k = 1 + xt.z;
  if (i < 1 || i > g.ni || j > g.nj) return;
  const uint32_t q = 4u * (uint32_t)g.at(i, j, k);
  const bool in = i <= g.ni - 1 && j <= g.nj - 1 && k <= g.nk - 1;
  // all 36 loads before the first store, so that none waits on a store
  float x[12], xf[12];
  bool ok[12];
  for (int f = 0; f < 12; f++) {
    ok[f] = in || (f == 0 && i == g.ni && j <= g.nj - 1 && k <= g.nk - 1) ||
            (f == 1 && j == g.nj && i <= g.ni - 1 && k <= g.nk - 1) ||
            (f == 2 && k == g.nk && i <= g.ni - 1 && j <= g.nj - 1);
    if (!ok[f]) continue;
    xf[f] = LD(a.xf[f], q);
    x[f] = m1fc2 * LD(a.x[f], q) + filcoe * (xf[f] + LD(a.xp[f], q));
  }
  for (int f = 0; f < 12; f++)
    if (ok[f] && (!in || f <= 3 || f == 7 || f == 9 || f == 10)) ST(a.x[f], q, x[f]);
  if (!in) return;
  const float pbr = MQ(CRESS_PBR, a.pbr, q, k), tb = MQ(CRESS_PTBR, a.ptbr, q, k);
  satadj_cell(thresq, pbr, x[3], tb, x[4], x[5], x[6], x[8], x[11]);
  ST(a.x[4], q, x[4]);
  ST(a.x[5], q, x[5]);
  ST(a.x[6], q, x[6]);
  ST(a.x[8], q, x[8]);
  ST(a.x[11], q, x[11]);
  satadj_cell(thresq, pbr, xf[3], tb, xf[4], xf[5], xf[6], xf[8], xf[11]);
  ST(a.xf[4], q, xf[4]);
  ST(a.xf[5], q, xf[5]);
  ST(a.xf[6], q, xf[6]);
  ST(a.xf[8], q, xf[8]);
  ST(a.xf[11], q, xf[11]);
}

__global__ void k_micro_seed(unsigned *p, unsigned v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
int micro_init_tables() {
  static int done = 0;
  if (done) return 0;
  const float c[41] = {
      0.0f, .7939e-10f, .7841e-9f, .3369e-8f, .4336e-8f, .5285e-8f, .3728e-8f, .1852e-8f, .2991e-9f,
      .4248e-9f, .7434e-9f, .1812e-8f, .4394e-8f, .9145e-8f, .1725e-7f, .3348e-7f, .1725e-7f, .9175e-8f,
      .4412e-8f, .2252e-8f, .9115e-9f, .4876e-9f, .3473e-9f, .4758e-9f, .6306e-9f, .8573e-9f, .7868e-9f,
      .7192e-9f, .6513e-9f, .5956e-9f, .5333e-9f, .4834e-9f, .4335e-9f, .3836e-9f, .3337e-9f, .2838e-9f,
      .2339e-9f, .1840e-9f, .1341e-9f, .8420e-10f, .3430e-10f};
  const float pk[41] = {
      0.0f, .4006f, .4831f, .5320f, .5307f, .5319f, .5249f, .4888f, .3894f, .4047f, .4318f, .4771f, .5183f,
      .5463f, .5651f, .5813f, .5655f, .5478f, .5203f, .4906f, .4447f, .4126f, .3960f, .4149f, .4320f, .4506f,
      .4483f, .4460f, .4433f, .4413f, .4382f, .4361f, .4340f, .4319f, .4298f, .4277f, .4256f, .4235f, .4214f,
      .4193f, .4172f};
  float ck[41];
  for (int t = 0; t < 41; t++) ck[t] = c[t] * expf(3.0f * pk[t] * mc::ln10);
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_ckoe), ck, sizeof(ck)) != hipSuccess) return CRESS_EHIP;
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_pkoe), pk, sizeof(pk)) != hipSuccess) return CRESS_EHIP;
  done = 1;
  return 0;
}

MicroK micro_fold(float dtb, float thresq, float dz) {
  using namespace mc;
  MicroK K;
  const float oned3 = 1.0f / 3.0f, oned6 = 1.0f / 6.0f;
  K.dtb = dtb;
  K.thresq = thresq;
  K.dz = dz;
  float cdv = 1.0f / t0;
  K.cnu = expf(.57013f * logf(nu0)) * cdv;
  K.cdv = expf(.55249f * logf(dv0)) * cdv;
  K.pdiaqr = .5f * (5.0f + bur) - 1.0f;
  K.pdiaqs = .5f * (5.0f + bus) - 1.0f;
  K.pdiaqg = .5f * (5.0f + bug) - 1.0f;
  K.cvntr = .31f * sqrtf(aur) * expf(oned3 * logf(sc)) * gf5bur;
  K.cvnts = .31f * sqrtf(aus) * expf(oned3 * logf(sc)) * gf5bus;
  K.cvntg = .31f * sqrtf(aug) * expf(oned3 * logf(sc)) * gf5bug;
  K.curq = oned6 * aur * gf4bur;
  K.cusq = oned6 * aus * gf4bus;
  K.cugq = oned6 * aug * gf4bug;
  K.rwdt2 = (100.0f / rhow) * dtb;
  K.cc45 = 4.e5f * cc;
  K.cknd = 2.0f * lmb20 * p20 / (t20 * diasl);
  K.cdar = boltz / (3.0f * cc * diasl);
  K.cclcr = .25f * aur * gf3bur * cc * dtb;
  K.cclcs = .25f * aus * gf3bus * cc * dtb;
  K.cclcg = .25f * aug * gf3bug * cc * dtb;
  K.cclri = .25f * aur * gf3bur * eir * cc * dtb;
  K.cclrs = ers * cc * cc * rhow * dtb;
  K.cclrg = erg * cc * cc * rhow * dtb;
  K.cclis = .25f * aus * gf3bus * eis * cc * dtb;
  K.cclig = .25f * aug * gf3bug * eig * cc * dtb;
  K.cclsr = esr * cc * cc * rhos * dtb;
  K.cclsg = esg * cc * cc * rhos * dtb;
  K.cclsg3 = cc * cc * rhos * dtb;
  K.cc2dtn = -2.0f * cc * dtb;
  K.cc2dt = 2.0f * cc * dtb;
  K.cagin_ag = r0 * expf(3.0f * logf(.5f * aui * eii * xl / rhoi * dtb));
  K.ccdtb4 = 2.0f * cc * dtb;
  K.qccm = oned6 * rhow * cc * diaqcm * diaqcm * diaqcm;
  K.diaqs0 = expf(oned3 * logf(6.0f * ms0 / (cc * rhos)));
  K.cagin_cv = r0 * expf(3.0f * logf(.5f * oned3 * aui * eii * xl / rhoi * dtb));
  K.ccncr = .104f * g * ecc * expf(-oned3 * logf(rhow)) * dtb;
  K.ccnsg = rhog * (1.0f / (rhog - rhos));
  K.cfrrg = 20.0f * 100.0f * cc * cc * rhow * dtb;
  K.mi0352 = .5f * 3.5e8f * mi0;
  K.mi0353 = oned3 * 3.5e8f * mi0;
  (void)ess;
  return K;
}

}  // namespace cress
