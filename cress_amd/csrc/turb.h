// The coefficients of the 1.5-order TKE closure (tubopt=2): the first four
// calls of s_turbdrv (turbdrv.f90:245-276) -- s_bruntv, s_defomten,
// s_defomssq, s_eddyvis -- for flat terrain without map factors
// (trnopt=0, mfcopt=0), no surface physics (sfcopt=0) and isotropic grids
// (isoopt=1).
#pragma once
#include "common.h"

namespace cress {

struct TurbArgs {
  int moist, cphopt;                  // fmois 'moist' / 'dry'; cloud physics on when cphopt != 0
  int wbc, ebc, sbc, nbc, bbc, tbc;
  int iw, ie, js, jn;                 // 4th-order reduction flags (0 on periodic sides)
  float dx, dy, dz, dtb, thresq;
  const float *jcb, *jcb8u, *jcb8v, *jcb8w, *pbr, *ptbr, *rbr;
  const float *pp, *ptp, *qv, *qall, *u, *v, *w, *tke;   // past level
  float *s11, *s22, *s33, *s12, *s13, *s23, *s31, *s32;  // deformation tensor
  float *ssq, *nsq8w, *rkh, *rkv, *priv;
  float *ufrc, *vfrc, *wfrc;                             // velocity forcings (turb_drv)
};

// Launches the kernels on the null stream, in the reference's order.
void turb_coef(const Grid &g, const TurbArgs &a);

// All of s_turbdrv (turbdrv.f90:245-287): turb_coef, then s_strsten
// (sfcopt=0) and s_turbuvw (trnopt=0, mfcopt=0), which overwrite ufrc,
// vfrc and wfrc over their interiors.
void turb_drv(const Grid &g, const TurbArgs &a);

// s_turbflx (turbflx.f90:72-74), trnopt=0, sfcopt=0.
void turb_flx(const Grid &g, float dx, float dy, float dz, const float *jcb, const float *s, const float *kh8u,
              const float *kh8v, const float *kv8w, float *h1, float *h2, float *h3);

// s_turbs (turbs.f90:72-75), trnopt=0, mfcopt=0.
void turb_s(const Grid &g, float dx, float dy, float dz, const float *jcb8u, const float *jcb8v, const float *h1,
            const float *h2, const float *h3, float *sfrc);

// s_kh8uv (kh8uv.f90:60), mfcopt=0.
void turb_kh8uv(const Grid &g, const float *rkh, float *rkh8u, float *rkh8v);

// The arrays and options of s_forcetke past its advection call.
struct TkeArgs {
  float dx, dy, dz;
  int vsp, ksp0;                       // vertical sponge on TKE; ksp0(2)
  const float *jcb, *jcb8u, *jcb8v, *jcb8w, *rst, *tkep, *priv, *nsq8w, *ssq;
  const float *rbct;                   // (1:ni,1:nj,1:nk,1:2), read when vsp
  float *rkh, *rkv, *tkefrc, *rkv8s;
};

// s_forcetke (forcetke.f90:225-351) after s_advs has written tkefrc, for
// tubopt=2, isoopt=1, trnopt=0, mfcopt=0, no smoothing or lateral sponge:
// shear production, dissipation, the flux divergence, the vertical sponge,
// the eddy diffusivities and the buoyancy term, on the null stream.
void turb_tke(const Grid &g, const TkeArgs &a);

}  // namespace cress
