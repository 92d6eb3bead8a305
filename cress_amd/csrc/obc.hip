// Open lateral boundaries (option 7) on gfx950: the radiative fills of one
// acoustic substep, the corner averages and the large-step scalar fills
// without microphysics.  Operation order follows the reference routines
// cited; the build's -ffp-contract=off keeps every product and sum rounded
// on its own, so results equal the reference's FP32 results bitwise.
#include "obc.h"

namespace cress {

#define AT(a, i, j, k) a[g.at(i, j, k)]

// rbcu (rbcu.f90:262-400), rbcv (rbcv.f90:262-400), rbcw (rbcw.f90:348-470)
// and rbcss (rbcss.f90:290-400) on the edges of their boxes.  One thread
// per (side position, level): blockIdx.y is k-1, and x runs over the four
// sides (west and east by j = 0..nj+1, south and north by i = 0..ni+1).
// The corner values these routines also form are overwritten by bc4news
// right after in the reference, so they are skipped here; no thread then
// reads a cell another thread writes.
__global__ void k_obc_ac_edges(Grid g, ObcAc c, ObcAcArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.y;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const int ly = nj + 2, lx = ni + 2;
  if (t >= 2 * ly + 2 * lx) return;
  const bool kuv = k >= 2 && k <= nk - 2, kw = k >= 2 && k <= nk - 1;
  if (t < 2 * ly) {
    const bool west = t < ly;
    const int j = west ? t : t - ly;
    if (west) {
      if (kuv && j >= 2 && j <= nj - 2) {
        const long p = g.at(1, j, k);
        a.unew[p] = a.uold[p] - c.cw * (a.uold[p + 1] - a.uold[p]);
        a.ppf[p] = a.ppf[p] - c.cw * (a.ppf[p + 1] - a.ppf[p]);
      }
      if (kuv && j >= 2 && j <= nj - 1) {
        const long p = g.at(1, j, k);
        a.vnew[p] = a.vold[p] - c.cw * (a.vold[p + 1] - a.vold[p]);
      }
      if (kw && j >= 2 && j <= nj - 2) {
        const long p = g.at(1, j, k);
        a.wf[p] = a.wf[p] - c.cw * (a.wf[p + 1] - a.wf[p]);
      }
    } else {
      if (kuv && j >= 2 && j <= nj - 2) {
        const long p = g.at(ni, j, k);
        a.unew[p] = a.uold[p] + c.ce * (a.uold[p - 1] - a.uold[p]);
        const long q = p - 1;   // p' at i = ni-1
        a.ppf[q] = a.ppf[q] + c.ce * (a.ppf[q - 1] - a.ppf[q]);
      }
      if (kuv && j >= 2 && j <= nj - 1) {
        const long p = g.at(ni - 1, j, k);
        a.vnew[p] = a.vold[p] + c.ce * (a.vold[p - 1] - a.vold[p]);
      }
      if (kw && j >= 2 && j <= nj - 2) {
        const long p = g.at(ni - 1, j, k);
        a.wf[p] = a.wf[p] + c.ce * (a.wf[p - 1] - a.wf[p]);
      }
    }
  } else {
    const int t2 = t - 2 * ly;
    const bool south = t2 < lx;
    const int i = south ? t2 : t2 - lx;
    const long sx = g.nx;
    if (south) {
      if (kuv && i >= 2 && i <= ni - 1) {
        const long p = g.at(i, 1, k);
        a.unew[p] = a.uold[p] - c.cs * (a.uold[p + sx] - a.uold[p]);
      }
      if (kuv && i >= 2 && i <= ni - 2) {
        const long p = g.at(i, 1, k);
        a.vnew[p] = a.vold[p] - c.cs * (a.vold[p + sx] - a.vold[p]);
        a.ppf[p] = a.ppf[p] - c.cs * (a.ppf[p + sx] - a.ppf[p]);
      }
      if (kw && i >= 2 && i <= ni - 2) {
        const long p = g.at(i, 1, k);
        a.wf[p] = a.wf[p] - c.cs * (a.wf[p + sx] - a.wf[p]);
      }
    } else {
      if (kuv && i >= 2 && i <= ni - 1) {
        const long p = g.at(i, nj - 1, k);
        a.unew[p] = a.uold[p] + c.cn * (a.uold[p - sx] - a.uold[p]);
      }
      if (kuv && i >= 2 && i <= ni - 2) {
        const long p = g.at(i, nj, k);
        a.vnew[p] = a.vold[p] + c.cn * (a.vold[p - sx] - a.vold[p]);
        const long q = p - sx;  // p' at j = nj-1
        a.ppf[q] = a.ppf[q] + c.cn * (a.ppf[q - sx] - a.ppf[q]);
      }
      if (kw && i >= 2 && i <= ni - 2) {
        const long p = g.at(i, nj - 1, k);
        a.wf[p] = a.wf[p] + c.cn * (a.wf[p - sx] - a.wf[p]);
      }
    }
  }
}

// bc4news (bc4news.f90:153-204): each field's box corners become the mean
// of their x and y neighbours, at every level 1..nk.  Thread (k, field).
__global__ void k_obc_corners(Grid g, CornerSet cs) {
  const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (k > g.nk || f >= cs.n) return;
  float *v = cs.ptr[f];
  const int iw = cs.isw[f], ie = cs.ise[f], js = cs.jss[f], jn = cs.jsn[f];
  AT(v, iw, js, k) = 0.5f * (AT(v, iw + 1, js, k) + AT(v, iw, js + 1, k));
  AT(v, ie, js, k) = 0.5f * (AT(v, ie - 1, js, k) + AT(v, ie, js + 1, k));
  AT(v, iw, jn, k) = 0.5f * (AT(v, iw + 1, jn, k) + AT(v, iw, jn - 1, k));
  AT(v, ie, jn, k) = 0.5f * (AT(v, ie - 1, jn, k) + AT(v, ie, jn - 1, k));
}

// steps.f90's fills of theta' (rbcpt) and qv (rbcqv) on the boundary ring
// of 1..ni-1 x 1..nj-1 at k=2..nk-2, corners from bc4news; the levels 1 and
// nk-1 follow from vbcs afterwards.  Thread (ring cell, k-2, scalar).
__global__ void k_obc_steps(Grid g, ObcLf lf, ObcStepsArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  const int sc = blockIdx.z;
  const int nxr = g.ni - 1, nyr = g.nj - 1;
  if (t >= 2 * nxr + 2 * (nyr - 2) || k > g.nk - 2) return;
  int i, j;
  if (t < nxr) {
    i = 1 + t;
    j = 1;
  } else if (t < 2 * nxr) {
    i = 1 + (t - nxr);
    j = nyr;
  } else if (t < 2 * nxr + (nyr - 2)) {
    i = 1;
    j = 2 + (t - 2 * nxr);
  } else {
    i = nxr;
    j = 2 + (t - 2 * nxr - (nyr - 2));
  }
  a.sf[sc][g.at(i, j, k)] = obc_lf_ring(g, lf, sc == 1, a.s[sc], a.sp[sc], i, j, k);
}

}  // namespace cress
