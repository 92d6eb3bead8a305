// Mass and energy budgets of one level of a context, reduced in binary64
// on the device: the conservation diagnostics of the step path (the only
// FP64 arithmetic on the GPU besides the base-state setup).  The reference
// monitors extremes (getmxn.f90:173-232, max/min over the domain gathered
// to the root) and accumulates the surface precipitation in upwqp
// (upwqp.f90:199-202: rate (F(1)+F(2))/(2 rhow), accumulation += rate dtp/2
// per sub-step of the leapfrog); these reductions add the totals a water
// and energy budget needs.  Reductions run in a fixed order (per-thread
// strided sums, a fixed tree per block, then one block over the partials),
// so a result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include "common.h"
#include "diag.h"

namespace cress {

namespace dg {
// comphy.f90
constexpr double g = 9.8, cp = 1004.0, rd = 287.0, epsva = 0.622, p0 = 1.e5, rhow = 1.e3;
constexpr double cv = cp - rd;
}  // namespace dg

constexpr int DIAG_BLOCK = 256;

// Partial sums of one block: sums[s] for the NSUM sums, then NEXT extremes.
__global__ void __launch_bounds__(DIAG_BLOCK) k_diag_cells(Grid g, DiagArgs a, double *part) {
  __shared__ double red[DIAG_BLOCK];
  const int tid = threadIdx.x;
  const long nci = g.ni - 3, ncj = g.nj - 3, nck = g.nk - 3;   // i,j,k = 2..n-2
  const long ncell = nci * ncj * nck;
  const long stride = (long)gridDim.x * DIAG_BLOCK;
  double s[DIAG_NSUM];
  double ex[DIAG_NEXT];
  for (int q = 0; q < DIAG_NSUM; q++) s[q] = 0.0;
  for (int q = 0; q < DIAG_NEXT; q++) ex[q] = (q % 2 == 0) ? -1.0e300 : 1.0e300;
  const double dv = a.dx * a.dy * a.dz;
  for (long e = (long)blockIdx.x * DIAG_BLOCK + tid; e < ncell; e += stride) {
    const int i = 2 + (int)(e % nci);
    const int j = 2 + (int)((e / nci) % ncj);
    const int k = 2 + (int)(e / (nci * ncj));
    const long p = g.at(i, j, k);
    const double rst = MP(CRESS_RST, a.rst, p, k), jcb = MP(CRESS_JCB, a.jcb, p, k);
    const double qv = a.q[p];
    const double ql = (double)a.q[p + a.qn] + (double)a.q[p + 2 * a.qn];
    const double qi = (double)a.q[p + 3 * a.qn] + (double)a.q[p + 4 * a.qn] + (double)a.q[p + 5 * a.qn];
    const double pr = (double)MP(CRESS_PBR, a.pbr, p, k) + (double)a.pp[p];
    const double th = (double)MP(CRESS_PTBR, a.ptbr, p, k) + (double)a.ptp[p];
    const double t = th * pow(pr / dg::p0, dg::rd / dg::cp);
    // dry-air density from the partial pressure of dry air, p (eps/(eps+qv))
    const double rhod = pr * dg::epsva / ((dg::epsva + qv) * dg::rd * t);
    const double md = jcb * rhod * dv;
    const double uc = 0.5 * ((double)a.u[p] + (double)a.u[p + 1]);
    const double vc = 0.5 * ((double)a.v[p] + (double)a.v[p + g.nx]);
    const double wc = 0.5 * ((double)a.w[p] + (double)a.w[p + g.nxy]);
    s[DIAG_QV] += rst * qv * dv;
    s[DIAG_QL] += rst * ql * dv;
    s[DIAG_QI] += rst * qi * dv;
    s[DIAG_DRY] += md;
    s[DIAG_KE] += 0.5 * md * (uc * uc + vc * vc + wc * wc);
    s[DIAG_IE] += md * dg::cv * t;
    s[DIAG_PE] += md * dg::g * a.zs[k];
    s[DIAG_THETA] += rst * (double)a.ptp[p] * dv;
    const double w = a.w[p], ptp = a.ptp[p];
    ex[DIAG_WMAX - DIAG_NSUM] = fmax(ex[DIAG_WMAX - DIAG_NSUM], w);
    ex[DIAG_WMIN - DIAG_NSUM] = fmin(ex[DIAG_WMIN - DIAG_NSUM], w);
    ex[DIAG_PTPMAX - DIAG_NSUM] = fmax(ex[DIAG_PTPMAX - DIAG_NSUM], ptp);
    ex[DIAG_PTPMIN - DIAG_NSUM] = fmin(ex[DIAG_PTPMIN - DIAG_NSUM], ptp);
  }
  // surface precipitation accumulated over the interior columns, as water mass
  const long ncol = nci * ncj;
  const long n2 = (long)(g.ni + 2) * (g.nj + 2);
  for (long e = (long)blockIdx.x * DIAG_BLOCK + tid; e < ncol && a.prwtr; e += stride) {
    const int i = 2 + (int)(e % nci), j = 2 + (int)(e / nci);
    const long c = (long)i + (long)(g.ni + 2) * j;
    double acc = 0.0;
    for (int sp = 0; sp < 2; sp++) acc += a.prwtr[c + n2 * (1 + 2 * sp)];
    for (int sp = 0; sp < 3 && a.price; sp++) acc += a.price[c + n2 * (1 + 2 * sp)];
    s[DIAG_PREC] += acc * dg::rhow * a.dx * a.dy;
  }
  double *out = part + (long)blockIdx.x * DIAG_N;
  for (int q = 0; q < DIAG_N; q++) {
    red[tid] = q < DIAG_NSUM ? s[q] : ex[q - DIAG_NSUM];
    __syncthreads();
    for (int h = DIAG_BLOCK / 2; h > 0; h >>= 1) {
      if (tid < h) {
        const double x = red[tid], y = red[tid + h];
        red[tid] = q < DIAG_NSUM ? x + y : ((q - DIAG_NSUM) % 2 == 0 ? fmax(x, y) : fmin(x, y));
      }
      __syncthreads();
    }
    if (tid == 0) out[q] = red[0];
    __syncthreads();
  }
}

// The partials of nb blocks, in block order.
__global__ void __launch_bounds__(DIAG_BLOCK) k_diag_final(const double *part, int nb, double *out) {
  const int q = threadIdx.x;
  if (q >= DIAG_N) return;
  double r = q < DIAG_NSUM ? 0.0 : ((q - DIAG_NSUM) % 2 == 0 ? -1.0e300 : 1.0e300);
  for (int b = 0; b < nb; b++) {
    const double x = part[(long)b * DIAG_N + q];
    r = q < DIAG_NSUM ? r + x : ((q - DIAG_NSUM) % 2 == 0 ? fmax(r, x) : fmin(r, x));
  }
  out[q] = r;
}

}  // namespace cress
