// Kernel argument blocks and launch entry points of the dynamics kernels.
#pragma once
#include "common.h"

namespace cress {

struct ForceIn {
  const float *u, *v, *w, *up, *vp, *wp, *ptp, *q, *qvbr, *ptbr, *ubr, *vbr;
  const float *rst, *rst8u, *rst8v, *rst8w, *rstxu, *rstxv, *rstxwc, *fc, *rbct;
  const float *jcb8u, *jcb8v, *jcb8w;  // for the smoothing of u, v, w
  long qn;  // species stride in q (qv,qc,qr,qi,qs,qg)
};

struct ScalarIn {
  const float *s[8];    // present level: ptp, qv, qc, qr, qi, qs, qg, nci
  const float *sp[8];   // past level
  float *frc[8];        // forcing outputs
  const float *rstxu, *rstxv, *rstxwc, *rst, *qvbr, *rbct, *ptbr, *rbr, *w;
};

// Launch shapes.  Each stencil and column kernel takes its rows per
// workgroup from blockDim.y (at most the _BYMAX below) and its tile order
// from Tiles, so one build holds every shape; the context's defaults (the
// fastest measured at 1024x1024x80) can be replaced per grid by the
// autotuner (cress_autotune).  Measurements at 1024x1024x80, ms per launch:
//   column kernel, metrics streamed as 3-D fields: 8 waves/SIMD 2.84, 5
//   (16 KiB LDS cap) 2.27, 3 (24 KiB) 2.33; metrics from per-level tables:
//   16 KiB 1.41, 12 KiB 1.27, 8 KiB 1.26, 4 KiB (registers decide) 1.26;
//   fused ptil/stepuv: x-fastest 64x8 1.405, 64x4 1.412; z-fastest 64x4
//   1.571, 64x8 1.580; from per-level tables 64x4 1.11, 64x8 1.14, 64x16
//   1.08, z-fastest 64x8 1.16; 64x4 threads each marching 16 rows in y
//   1.17-1.21; own-load-first 64x16 0.99, with the slot layout 64x8 0.97;
//   scalar forcing 64x4 3.59, 64x8 3.79, 64x16 3.66, 64x16 x-fastest 4.97;
//   u/v forcing 64x4 1.91, 64x8 2.01, 64x16 2.42.
constexpr int CRESS_PTUV_BYMAX = 16;
constexpr int CRESS_COL_BYMAX = 4;
constexpr int CRESS_FS_BYMAX = 8;

// The tunable launch parameters of one context (defaults and ranges in
// capi.hip: tune_default, cress_autotune).
struct LaunchTune {
  int ptuv_by, ptuv_slots, ptuv_order;   // fused ptil/stepuv
  int col_by, col_lds;                   // column kernel (LDS bytes cap occupancy)
  int fs_by, fs_order;                   // scalar forcing
  int ff_by, ff_order;                   // u/v, w and p' forcing
};

struct AcPtUv {
  float dts, divcv;
  const float *jcb, *rst8u, *rst8v, *rst8w, *uf, *vf, *wc, *ppf, *ufrc, *vfrc;
  float *ptil, *ufn, *vfn;   // ufn/vfn: the other buffer of the uf/vf pair
};

struct AcCol {
  float dts;
  const float *uf, *vf, *pfrc, *wfrc;
  float *ptil;  // read; its consumed slots hold the Thomas super-diagonal
  const float *jcb, *jcb8u, *jcb8v, *jcb8w, *rcsq, *rst, *rbr, *rst8w;
  float *wf, *wc, *ppf;
  // With rst8w, rcsq, rst, rbr, jcb, jcb8u, jcb8v, jcb8w horizontally uniform
  // (per-level tables, read without a run-time test), the Thomas
  // coefficients of stepwi/gaussel depend on the level and dts only:
  // coef[3k] = rr(k), coef[3k+1] = 1/(ss(k) - rr(k) ss'(k-1)) (1/ss(3) at
  // k=3), coef[3k+2] = ss'(k), k = 3..nk-2 (k_ac_col<true>); else null.
  const float *coef;
};

struct StepsIn {
  const float *ptpp, *ptfrc, *qp, *qvfrc, *ncip;
  float *ptpf, *qf, *ncif;
  long qn;
};

struct FillSpec {
  float *ptr;
  int iwsnd, iwrcv, iesnd, iercv, jssnd, jsrcv, jnsnd, jnrcv, kmax;
};
struct FillSet {
  int n;
  FillSpec f[16];
};

struct VbcSet {
  int n;
  int kind[16];  // 0 scalar, 1 u, 2 v, 3 p
  float *ptr[16];
};

struct AsselinSet {
  int n;
  int kind[12];  // 0 scalar, 1 u, 2 v, 3 w
  float *x[12];
  const float *xp[12];
  const float *xf[12];
};

// U: every base/metric field the kernel reads is a per-level table (Grid::
// umask has all of CRESS_JCB..CRESS_RCSQ), read without a run-time test.
template <bool U>
__global__ void k_phy2cnt(Tiles T, Grid g, Opts o, const float *w, const float *jcb8w, float *wc);
template <bool U>
__global__ void k_prep(Tiles T, Grid g, Opts o, const float *u, const float *v, const float *w,
                       const float *rst8u, const float *rst8v, const float *rst8w,
                       const float *jcb8w, float *wc, float *rstxu, float *rstxv, float *rstxwc);
// SM: with the 4th-order smoothing (smtopt 2/3) compiled in
template <bool SM, bool U>
__global__ void k_force_uv(Tiles T, Grid g, Opts o, Consts c, ForceIn in, float *ufrc, float *vfrc);
template <bool SM, bool U>
__global__ void k_force_w(Tiles T, Grid g, Opts o, Consts c, ForceIn in, float *wfrc);
template <bool U>
__global__ void k_force_p(Tiles T, Grid g, Opts o, Consts c, const float *jcb, const float *jcb8u,
                          const float *jcb8v, const float *jcb8w, const float *u, const float *v,
                          const float *wc, const float *pp, const float *ppp, const float *rbct,
                          float *pfrc);
template <bool SM, bool U>
__global__ void k_force_s(Tiles T, Grid g, Opts o, Consts c, ScalarIn in, int nscalar);
// U: rst8u, rst8v, rst8w and jcb are per-level tables; SLOTS: the ptil
// tile and its halo in slots of consecutive points (else one point per
// thread, the halo row and column by the first row and lane)
template <bool U, int SLOTS>
__global__ void k_ac_ptuv(Tiles T, Grid g, Opts o, Consts c, AcPtUv a);
template <bool UNI>
__global__ void k_ac_col(Tiles T, Grid g, Opts o, Consts c, AcCol a);
__global__ void k_ac_ring(Grid g, Opts o, const float *jcb8w, float *wf, float *wc, float *ppf);
__global__ void k_steps(Tiles T, Grid g, float dtb2, const float *rst, StepsIn s, int withq);
__global__ void k_bcycle_x(Grid g, FillSet fs);
__global__ void k_bcycle_y(Grid g, FillSet fs);
__global__ void k_vbc(Grid g, int bbc, VbcSet vs);
__global__ void k_timeflt(Tiles T, Grid g, Consts c, AsselinSet as);

}  // namespace cress

namespace cress {
__global__ void k_gaussel(Grid g, int istr, int iend, int jstr, int jend, int kstr, int kend,
                          const float *rr, float *ss, const float *tt, float *ff);
__global__ void k_advs_only(Grid g, Opts o, Consts c, ScalarIn in);
__global__ void k_advuvw_only(Grid g, Opts o, Consts c, const float *u, const float *v, const float *w,
                              const float *rxu, const float *rxv, const float *rxw, float *ufrc, float *vfrc,
                              float *wfrc);
__global__ void k_bc8(Grid g, int kind, int lo, int hi, int kmax, float *var);
__global__ void k_dtdrst(Grid g, float dtb2, const float *rst, float *dtdrst);
__global__ void k_steps_tke(Grid g, float dtb2, const float *rst, const float *tkep, float *tkef);
}  // namespace cress
