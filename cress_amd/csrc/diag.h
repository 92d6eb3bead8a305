// FP64 mass and energy diagnostics of a context level (diag.hip).
#pragma once
#include "common.h"

namespace cress {

// Result slots (the C ABI's cress_diag order): sums first, then extremes
// (even: maximum, odd: minimum).
enum {
  DIAG_QV = 0, DIAG_QL, DIAG_QI, DIAG_PREC, DIAG_DRY, DIAG_KE, DIAG_IE, DIAG_PE, DIAG_THETA,
  DIAG_NSUM,
  DIAG_WMAX = DIAG_NSUM, DIAG_WMIN, DIAG_PTPMAX, DIAG_PTPMIN,
  DIAG_N
};
constexpr int DIAG_NEXT = DIAG_N - DIAG_NSUM;

struct DiagArgs {
  const float *u, *v, *w, *pp, *ptp, *q;   // q: the six species blocks
  long qn;
  const float *rst, *jcb, *pbr, *ptbr;     // base state (or per-level tables)
  const float *prwtr, *price;              // (0:ni+1,0:nj+1,2,2|3), may be null
  const double *zs;                        // height of the scalar levels, [k]
  double dx, dy, dz;
};

__global__ void k_diag_cells(Grid g, DiagArgs a, double *part);
__global__ void k_diag_final(const double *part, int nb, double *out);

}  // namespace cress
