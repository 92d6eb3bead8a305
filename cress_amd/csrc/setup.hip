// Base state and metrics of a context from a 1-D sounding, in binary64 on
// the device (the setup is, with the conservation diagnostics, the only
// binary64 arithmetic of the GPU path).  Flat terrain: every column is the
// same, so one lane forms the per-level values in the reference's order of
// routines and operations, and a second kernel broadcasts them, rounded once
// to binary32, into the 3-D fields the large step reads.
//
//   s_stretch   (stretch.f90:174-359)   zeta-level heights zsth, sthopt 0/1
//   s_jacobian  (jacobian.f90:243-251)  jcb = (zph(k+1)-zph(k))/(z(k+1)-z(k))
//   var8w8s     (setbase.f90:153-160)   scalar-level heights zph8s
//   s_intrpsnd  (intrpsnd.f90:99-301)   the sounding at zph8s by vint13's
//               linear interpolation / extrapolation (vint13.f90)
//   s_setbase   (setbase.f90:171-185)   ptvbr, pibr, rbr
//   s_bcbase    (bcbase.f90:120-190)    levels 1 and nk-1
//   s_baserho   (baserho.f90:153-170)   rst = |jcb| rbr
//   s_var8uvw   (var8uvw.f90:166-294)   w-point averages, bc8w (bbc, tbc)
//   s_sndwave   (sndwave.f90:119)       rcsq = (cp/cv) pbr
//   s_vspdmp    (vspdmp.f90:269-285)    the Rayleigh sponge, vspopt=2
#include <hip/hip_runtime.h>

#include "setup.h"

namespace cress {

namespace su {
// comphy.f90
constexpr double g = 9.8, cp = 1004.0, cv = 717.0, rd = 287.0, p0 = 1.e5, epsav = 1.6077;
constexpr double cc = 3.141592;   // commath.f90:42
}  // namespace su

// vint13: linear interpolation of a 1-D profile at height z, linear
// extrapolation from the end intervals outside it.
__device__ double vint13(const SetupArgs &a, const double *var, double z) {
  const int n = a.nlev;
  const double *zz = a.z;
  if (zz[0] > z) {
    const double dk = (z - zz[0]) / (zz[1] - zz[0]);
    return (1.0 - dk) * var[0] + dk * var[1];
  }
  if (zz[n] <= z) {
    const double dk = (z - zz[n - 1]) / (zz[n] - zz[n - 1]);
    return (1.0 - dk) * var[n - 1] + dk * var[n];
  }
  for (int kl = 0; kl < n; kl++)
    if (zz[kl] <= z && zz[kl + 1] > z) {
      const double dk = (z - zz[kl]) / (zz[kl + 1] - zz[kl]);
      return (1.0 - dk) * var[kl] + dk * var[kl + 1];
    }
  return var[n];
}

// bc8w for a w-point average var8w(2..nk-1): rigid (2) copies across the
// wall, zero-gradient (3) from the next level (bc8w.f90:130-184).
__device__ void bc8w_col(double *v, int nk, int bbc, int tbc) {
  v[1] = bbc == 2 ? v[3] : v[2];
  v[nk] = tbc == 2 ? v[nk - 2] : v[nk - 1];
}

// The per-level columns, index k = 1..nk (slot 0 unused), into col[c*(nk+1)+k].
__global__ void k_setup_col(SetupArgs a, double *col) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nk = a.nk;
  const int ld = nk + 1;
  double *zph = col + SC_ZPH * ld, *z8s = col + SC_Z8S * ld;
  double *ubr = col + SC_UBR * ld, *vbr = col + SC_VBR * ld, *pbr = col + SC_PBR * ld;
  double *ptbr = col + SC_PTBR * ld, *qvbr = col + SC_QVBR * ld, *rbr = col + SC_RBR * ld;
  double *jcb = col + SC_JCB * ld, *jcb8w = col + SC_JCB8W * ld, *rst = col + SC_RST * ld;
  double *rst8w = col + SC_RST8W * ld, *rcsq = col + SC_RCSQ * ld, *rbct = col + SC_RBCT * ld;
  double *ptv = col + SC_PTVBR * ld, *pib = col + SC_PIBR * ld;
  for (int c = 0; c < SC_N; c++)
    for (int k = 0; k <= nk; k++) col[c * ld + k] = 0.0;
  // s_stretch: zsth at the zeta levels; zph = zsth over flat terrain
  const double dz = a.dz, zsfc = a.zsfc;
  const double ztop2 = zsfc + (double)(nk - 3) * dz;
  double layer1 = a.layer1, layer2 = a.layer2;
  layer1 = fmax(fmin(ztop2 - zsfc, layer1), 0.0);
  double z1 = zsfc + layer1;
  layer2 = fmax(fmin(ztop2 - z1, layer2), 0.0);
  double z2 = z1 + layer2;
  const double dzmin = a.dzmin;
  if (a.sthopt == 0) {
    for (int k = 1; k <= nk; k++) zph[k] = zsfc + (double)(k - 2) * dz;
  } else {
    z1 = fmax(zsfc, z1);
    z2 = fmin(ztop2, z2);
    const int klow = (int)((z1 - zsfc) / dzmin + 0.1);
    double dzmid = 0.0, dzhigh = 0.0;
    int kmid = 0;
    if (z2 >= ztop2) {
      dzmid = (ztop2 - zsfc - (double)klow * dzmin) / (double)(nk - 3 - klow);
      dzhigh = 2.0 * dzmid - dzmin;
      kmid = 0;
    } else {
      double aa = 2.0 * (double)(nk - 3 - klow);
      const double bb = 2.0 * zsfc - ztop2 - z2 - (double)(nk - 3 - 3 * klow) * dzmin;
      const double cc = dzmin * (z2 - zsfc - (double)klow * dzmin);
      dzmid = 0.5 * (sqrt(bb * bb - 4.0 * aa * cc) - bb) / aa;
      kmid = (int)((ztop2 - z2) / (2.0 * dzmid - dzmin) + 0.1);
      aa = zsfc + (double)(klow - kmid) * dzmin + (double)(nk - 3 - klow + kmid) * dzmid;
      if (kmid != 0) {
        dzhigh = (2.0 * dzmid - dzmin) + (ztop2 - aa) / (double)kmid;
      } else if (nk - 3 - klow - kmid != 0) {
        dzmid = dzmid + (ztop2 - aa) / (double)(nk - 3 - klow - kmid);
        dzhigh = 2.0 * dzmid - dzmin;
      }
    }
    const double sa = dzmid - dzmin;
    const double sb = 1.0 / (double)(nk - 4 - kmid - klow);
    // dzsth(k): dzmin up to klow+1, the cubic in between, dzhigh above
    auto dzs = [&](int k) {
      if (k <= klow + 1) return dzmin;
      if (k <= nk - 2 - kmid) {
        const double t = 2.0 * sb * (double)(k - 2 - klow) - 1.0;
        return dzmid + sa * t * t * t;
      }
      return dzhigh;
    };
    zph[2] = zsfc;
    for (int k = 3; k <= nk - 2; k++) zph[k] = zph[k - 1] + dzs(k - 1);
    zph[1] = zph[2] - dzs(1);
    zph[nk - 1] = ztop2;
    zph[nk] = zph[nk - 1] + dzs(nk - 2);
  }
  // s_jacobian: the zeta spacing z(k+1) - z(k) is dz on every level
  for (int k = 1; k <= nk - 1; k++) jcb[k] = (zph[k + 1] - zph[k]) / dz;
  jcb[nk] = jcb[nk - 1];
  for (int k = 1; k <= nk - 1; k++) z8s[k] = 0.5 * (zph[k + 1] + zph[k]);
  z8s[nk] = z8s[nk - 1];
  // s_intrpsnd (vint13) for k = 2..nk-2, then s_setbase
  for (int k = 2; k <= nk - 2; k++) {
    ubr[k] = vint13(a, a.u, z8s[k]);
    vbr[k] = vint13(a, a.v, z8s[k]);
    pbr[k] = vint13(a, a.p, z8s[k]);
    ptbr[k] = vint13(a, a.pt, z8s[k]);
    qvbr[k] = vint13(a, a.qv, z8s[k]);
    ptv[k] = ptbr[k] * (1.0 + su::epsav * qvbr[k]) / (1.0 + qvbr[k]);
    pib[k] = exp((su::rd / su::cp) * log((1.0 / su::p0) * pbr[k]));
    rbr[k] = pbr[k] / (su::rd * ptv[k] * pib[k]);
  }
  // s_bcbase: levels 1 and nk-1
  ubr[1] = ubr[2];
  ubr[nk - 1] = ubr[nk - 2];
  vbr[1] = vbr[2];
  vbr[nk - 1] = vbr[nk - 2];
  ptbr[1] = ptbr[2];
  ptbr[nk - 1] = ptbr[nk - 2];
  qvbr[1] = qvbr[2];
  qvbr[nk - 1] = qvbr[nk - 2];
  ptv[1] = ptv[2];
  ptv[nk - 1] = ptv[nk - 2];
  const double gdvcp2 = 2.0 * su::g / su::cp, cpdvrd = su::cp / su::rd;
  pib[1] = pib[2] + gdvcp2 * (z8s[2] - z8s[1]) / (ptv[1] + ptv[2]);
  pib[nk - 1] = pib[nk - 2] - gdvcp2 * (z8s[nk - 1] - z8s[nk - 2]) / (ptv[nk - 2] + ptv[nk - 1]);
  pbr[1] = su::p0 * exp(cpdvrd * log(pib[1]));
  pbr[nk - 1] = su::p0 * exp(cpdvrd * log(pib[nk - 1]));
  rbr[1] = pbr[1] / (su::rd * ptv[1] * pib[1]);
  rbr[nk - 1] = pbr[nk - 1] / (su::rd * ptv[nk - 1] * pib[nk - 1]);
  // the top slot nk of the scalar columns, which no routine sets, repeats nk-1
  ubr[nk] = ubr[nk - 1];
  vbr[nk] = vbr[nk - 1];
  pbr[nk] = pbr[nk - 1];
  ptbr[nk] = ptbr[nk - 1];
  qvbr[nk] = qvbr[nk - 1];
  rbr[nk] = rbr[nk - 1];
  // s_baserho and s_var8uvw (w points) with bc8w; jcb8w likewise
  for (int k = 1; k <= nk - 1; k++) rst[k] = fabs(jcb[k]) * rbr[k];
  rst[nk] = rst[nk - 1];
  for (int k = 2; k <= nk - 1; k++) {
    rst8w[k] = 0.5 * (rst[k - 1] + rst[k]);
    jcb8w[k] = 0.5 * (jcb[k - 1] + jcb[k]);
  }
  bc8w_col(rst8w, nk, a.bbc, a.tbc);
  bc8w_col(jcb8w, nk, a.bbc, a.tbc);
  // s_sndwave
  for (int k = 1; k <= nk; k++) rcsq[k] = (su::cp / su::cv) * pbr[k];
  // s_vspdmp, vspopt=2: at the zeta levels above botbar
  const double cbar05 = 0.5 * a.vspbar;
  for (int k = 2; k <= nk - 1; k++)
    rbct[k] = zph[k] > a.botbar ? cbar05 * (1.0 - cos(su::cc * (zph[k] - a.botbar) / (zph[nk - 1] - a.botbar)))
                                : 0.0;
}

// The 3-D base fields (field ids CRESS_JCB..CRESS_RCSQ, x fastest, padded
// rows), the per-level tables, the sponge (1:ni,1:nj,1:nk,2) and the
// Coriolis array, from the columns, each value rounded once to binary32.
__global__ void k_setup_fill(Grid g, SetupFill f) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nk = g.nk, ld = nk + 1;
  const long nxy = g.nxy;
  if (p >= nxy * nk) return;
  const int k = 1 + (int)(p / nxy);
  const double *c = f.col;
  const float jcb = (float)c[SC_JCB * ld + k];
  const float v[CRESS_NBASE] = {jcb,
                                jcb,
                                jcb,
                                (float)c[SC_JCB8W * ld + k],
                                (float)c[SC_UBR * ld + k],
                                (float)c[SC_VBR * ld + k],
                                (float)c[SC_PBR * ld + k],
                                (float)c[SC_PTBR * ld + k],
                                (float)c[SC_QVBR * ld + k],
                                (float)c[SC_RBR * ld + k],
                                (float)c[SC_RST * ld + k],
                                (float)c[SC_RST * ld + k],
                                (float)c[SC_RST * ld + k],
                                (float)c[SC_RST8W * ld + k],
                                (float)c[SC_RCSQ * ld + k]};
  for (int id = 0; id < CRESS_NBASE; id++) f.base[id][p] = v[id];
  if (p % nxy == 0)
    for (int id = 0; id < CRESS_NBASE; id++) f.tab[id * nk + k - 1] = v[id];
  // sponge and Coriolis over their own layouts
  const long nr = (long)g.ni * g.nj * nk;
  for (long q = p; q < 2 * nr; q += nxy * nk) {
    const int kk = 1 + (int)((q % nr) / ((long)g.ni * g.nj));
    f.rbct[q] = q < nr ? 0.0f : (float)c[SC_RBCT * ld + kk];
  }
  const long n2 = (long)(g.ni + 2) * (g.nj + 2);
  for (long q = p; q < 2 * n2; q += nxy * nk) f.fc[q] = (float)(0.25 * f.fcor);
}

}  // namespace cress
