// Base state and metrics from a 1-D sounding in binary64 on the device
// (setup.hip, cress_set_base).
#pragma once
#include "common.h"

namespace cress {

// Per-level columns of the setup (k = 1..nk), cress_base_columns order.
enum {
  SC_ZPH = 0, SC_Z8S, SC_UBR, SC_VBR, SC_PBR, SC_PTBR, SC_QVBR, SC_RBR, SC_JCB, SC_JCB8W, SC_RST, SC_RST8W,
  SC_RCSQ, SC_RBCT, SC_PTVBR, SC_PIBR, SC_N
};

struct SetupArgs {
  int nk, sthopt, bbc, tbc;
  double dz, zsfc, dzmin, layer1, layer2;   // &gridsth
  double botbar, vspbar;                    // vspopt=2 sponge
  int nlev;                                 // the sounding: levels 0..nlev
  const double *z, *u, *v, *p, *pt, *qv;
};

struct SetupFill {
  const double *col;
  float *base[CRESS_NBASE];
  float *tab;
  float *rbct, *fc;
  double fcor;
};

__global__ void k_setup_col(SetupArgs a, double *col);
__global__ void k_setup_fill(Grid g, SetupFill f);

}  // namespace cress
