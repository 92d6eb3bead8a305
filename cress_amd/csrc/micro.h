// Bulk cold-rain microphysics (cphopt=2, haiopt=0): the kernels of
// s_cloudphy (cloudphy.f90:277-470 -> coldblk.f90:321-481) and s_satadjst
// (satadjst.f90:148-189).  The context's step driver launches them.
#pragma once
#include "common.h"
#include "mlp.h"
#include "obc.h"

namespace cress {

struct MicroScratch {
  enum { NF = 3 };
  float *f[NF];   // fall speeds of rain, snow, graupel (termblk) for the sedimentation kernel
  float *col;     // one word: binary32 bits of the sedimentation CFL minimum
};

// Per-call constants, folded on the host in binary32 in the reference's
// operation order (each routine's local parameters, dtb = microphysics step).
struct MicroK {
  float dtb, thresq, dz;
  float cnu, cdv, pdiaqr, pdiaqs, pdiaqg, cvntr, cvnts, cvntg;    // setblk.f90:280-330
  float curq, cusq, cugq;                                         // termblk.f90:280-300
  float rwdt2, cc45, cknd, cdar;                                  // nuc1stc.f90:172-200
  float cclcr, cclcs, cclcg, cclri, cclrs, cclrg, cclis, cclig,   // collect.f90:297-321
      cclsr, cclsg, cclsg3;
  float cc2dtn, cc2dt, cagin_ag, ccdtb4, qccm, diaqs0, cagin_cv,  // prodctwg, melting, aggregat,
      ccncr, ccnsg, cfrrg, mi0352, mi0353;                        // depsit, convers, freezing, nuc2nd
};

struct MicroArgs {
  const float *jcb, *pbr, *ptbr, *rbr, *rst;
  const float *ptfrc, *qvfrc, *sfrc;    // steps forcings (sfrc: qc, qr, qi, qs, qg, nci)
  // past level: qp holds qv, qpw the liquid species (qc, qr) and qpi the
  // ice species (qi, qs, qg), each block of qn floats per species as the
  // reference's qv, qwtr and qice arrays; the context's stacked six-species
  // block sets qpw = qp + qn, qpi = qp + 3 qn
  const float *ppp, *ptpp, *qp, *qpw, *qpi, *ncip;
  float *ptpf, *qf, *qfw, *qfi, *ncif;  // future level, the same layout
  float *urq, *usq, *ugq;               // fall speeds handed to k_micro_fall
  unsigned *dtpmin;                     // binary32 bits of the CFL minimum
  long qn;                              // floats per species block
  // open lateral boundaries: the present level the radiative fills read and
  // their gammas; open=0 on a periodic domain
  const float *ptpc, *qc, *ncic;
  ObcLf lf;
  int open;
  // 1: the future level already holds what steps leaves (s_cloudphy's
  // input on its own); 0: it is formed here from the past level and the
  // forcings
  int given;
  // sources: 0 srcblk; 1 srcblk and record (state, tendency) pairs into
  // cap; 2 per cell as k_mlp_select chose (mflag), the network's
  // tendencies from mout (7 blocks of qn floats)
  int mode;
  MlpCapture cap;
  const int *mflag;
  const float *mout;
};

// The network's choice per cell and the list of its cells (MlpStep), from
// the past level: no active process (no hydrometeor above thresq and no
// ice nucleation from vapour) -> 2; active inside the training envelope ->
// 1, with the normalised features appended to the list; else 0 (srcblk).
struct MlpSelectArgs {
  const float *pbr, *ptbr, *rbr, *ppp, *ptpp, *qp, *qpw, *qpi;
  long qn;
  float thresq;
  MlpStep s;
  const float *norm;
};
__global__ void k_mlp_select(Tiles T, Grid g, MlpSelectArgs a);

// First column of the sedimentation kernel's x tiles (measured at
// 1024x1024x80: from i=1 2.85 ms, from i=0 3.24 ms per launch).
#ifndef CRESS_FALL_ORG
#define CRESS_FALL_ORG 1
#endif

struct FallArgs {
  const float *rbr, *rst, *qpw, *qpi, *urq, *usq, *ugq;  // species blocks as in MicroArgs
  float *qfw, *qfi, *ncif, *prwtr, *price;
  const unsigned *dtpmin;
  long qn;   // floats per species block
  long n2;   // floats per 2-D level of the precipitation arrays
  float dtb, dziv, thresq;
};

// The twelve prognostics at three levels (u, v, w, pp, ptp, qv, qc, qr, qi,
// qs, qg, nci) and the base state the saturation adjustment needs.
struct AsselinSat {
  float *x[12];
  const float *xp[12];
  float *xf[12];
  const float *pbr, *ptbr;
};

// Koenig table into constant memory (once per process and device set).
int micro_init_tables();
MicroK micro_fold(float dtb_micro, float thresq, float dz);

__global__ void k_micro_seed(unsigned *p, unsigned v);
__global__ void k_micro_src(Tiles T, Grid g, MicroK K, MicroArgs a);
__global__ void k_micro_fall(Tiles T, Grid g, FallArgs a);
__global__ void k_asselin_sat(Tiles T, Grid g, float m1fc2, float filcoe, float thresq, AsselinSat a);
__global__ void k_satadjst(Tiles T, Grid g, float thresq, const float *pbr, const float *ptbr, const float *pp, float *ptp,
                           float *q, float *nci, long qn);

}  // namespace cress
