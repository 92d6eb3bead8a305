// Bulk cold-rain microphysics (cphopt=2) entry points: s_cloudphy
// (cloudphy.f90:277-470) and s_satadjst (satadjst.f90:148-189).
#pragma once
#include "common.h"

namespace cress {

struct MicroScratch {
  enum { NF = 16 };
  float *f[NF];   // 3-D work fields
  float *col;     // small reduction buffer (fall-speed CFL minimum)
};

int micro_cloudphy(const Grid &g, const Consts &c, const Opts &o, hipStream_t st, float dtb, int first,
                   float *const *base, const float *ptpp, const float *qp, const float *ncip, const float *ppp,
                   float *ptpf, float *qf, float *ncif, const float *wf, float *prwtr, float *price,
                   MicroScratch *ms);
int micro_satadjst(const Grid &g, const Consts &c, const Opts &o, hipStream_t st, float *const *base,
                   const float *pp, float *ptp, float *q, float *nci, MicroScratch *ms);

}  // namespace cress
