/* CPU restatement of one CReSS large time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is the oracle the HIP path is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Compiled twice from cress_oracle.c: REAL=float  -> liboracle32.so,
 *                                     REAL=double -> liboracle64.so.
 * Each routine restates one reference routine (file:line cited at the
 * definition) with the same loop ranges and the same floating-point operation
 * order, so that the FP32 build reproduces the reference's FP32 arithmetic.
 *
 * Arrays: Fortran real(0:ni+1,0:nj+1,1:nk), column-major, x fastest.
 */
#ifndef CRESS_ORACLE_H
#define CRESS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifndef REAL
#define REAL float
#endif

typedef struct {
  int ni, nj, nk;
  int wbc, ebc, sbc, nbc, bbc, tbc;   /* lateral -1 = periodic at one rank */
  int advopt, impopt, gwmopt, buyopt, divopt, cphopt, haiopt;
  int vspopt, coropt, sthopt;
  int iwest, ieast, jsouth, jnorth;   /* 4th-order reduction flags */
  int ksp0[2];
  char vspvar[10];                    /* 'o'/'x' per u,v,w,p,pt,qv,q,... */
  double dx, dy, dz, weicoe, filcoe, thresq;
  double gwave;                       /* open-boundary phase speed (m/s) */
  int smtopt;                         /* 0 none, 2/3 4th-order smoothing */
  double smhcoe, smvcoe;              /* smoothing coefficients (1/s) */
} or_cfg;

#endif
