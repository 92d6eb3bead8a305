/* CPU restatement of one CReSS large time step -- TEST INFRASTRUCTURE ONLY.
 * See cress_oracle.h.  Loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, always as the checker, never as the product.
 *
 * Conventions shared by every routine:
 *   - arrays are real(0:ni+1,0:nj+1,1:nk) (Fortran column-major, x fastest);
 *   - loop ranges and the order of floating-point operations follow the
 *     reference routine cited at each definition, including its explicit
 *     parentheses and Fortran's left-to-right evaluation, so that the FP32
 *     build (OR_FP32) reproduces the reference's FP32 results;
 *   - option branches not taken by the supported configuration (flat terrain
 *     trnopt=0, mfcopt=0, tubopt=0, smtopt=0, exbopt=0, single rank) are not
 *     restated; cfg_check() rejects configurations outside that set.
 *
 * Pinning (tests/): bitwise against the reference's own compiled routines
 * for advs, advuvw, rstuvwc, buoywb, coriuv, vspuvw, forcep/advp,
 * forcept/forceqv/forceq, pgrad/diver3d, diverpih, diverpiv, advbspi,
 * pgradiv, buoywsi, gaussel, bcycle, vbc*, totalqwi, timeflt and the whole
 * microphysics of cress_micro.c; on open sides (option 7, cress_obc.c)
 * also phvuvw, phvs, rbcu, rbcv, rbcw, rbcss, rbcpt, rbcqv, rbcq, rbcs,
 * bc4news, the advection routines with the wall reduction, and the lbc*
 * routines' being no-ops.  stepuv, stepwi, steppi, steps, the
 * drivers and shift2nd cannot be built from the reference (their modules
 * reach MPI); they are pinned by whole-step invariants (mirror, transpose,
 * translation, boundedness).  A mistake in them that is symmetric and
 * stable (e.g. a wrong but symmetric coefficient of stepwi's tridiagonal
 * system) is not caught: for those routines parity is unpinned against
 * the reference.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "cress_oracle.h"

#ifdef OR_FP32
#define K(x) x##f
#define EXP expf
#define LOG logf
#define SQRT sqrtf
#define POW powf
#define COS cosf
#else
#define K(x) x
#define EXP exp
#define LOG log
#define SQRT sqrt
#define POW pow
#define COS cos
#endif

typedef REAL R;

/* OpenMP threads the parallel sweeps below use */
int or_threads(void) { return omp_get_max_threads(); }

/* k-sweeps whose iterations write only level k run across OpenMP threads
 * (the per-cell arithmetic, hence the result, is the same for any thread
 * count); recurrences in k and reductions stay serial. */
#define OMP_FOR _Pragma("omp parallel for schedule(static)")

#define DECL_DIMS(c) \
  const int ni = (c)->ni, nj = (c)->nj, nk = (c)->nk; \
  const size_t NX = (size_t)ni + 2, NY = (size_t)nj + 2; \
  (void)ni; (void)nj; (void)nk; (void)NX; (void)NY
#define A3(a, i, j, k) (a)[(size_t)(i) + NX * ((size_t)(j) + NY * ((size_t)(k) - 1))]

/* commath.f90:63,75 and comphy.f90:49 -- folded in REAL like the reference */
#define ONED24 ((R)1 / (R)24)
#define FOURD3 ((R)4 / (R)3)
#define GRAV K(9.8)

static int nfield(const or_cfg *c) { return (c->ni + 2) * (c->nj + 2) * c->nk; }

int or_real_size(void) { return (int)sizeof(R); }

/* ------------------------------------------------------------------------ */
/* Boundary / halo fills                                                    */
/* ------------------------------------------------------------------------ */

/* s_bc8u (bc8u.f90:132-208), s_bc8v (bc8v.f90:132-208), s_bc8w
 * (bc8w.f90:130-188) on one rank whose sides are all outer boundaries
 * (ebw=ebe=ebs=ebn=1): kind 0 u points, x sides over j=0..nj+1,
 * k=1..kmax, west then east; kind 1 v points, y sides over i=0..ni+1;
 * kind 2 w points, bottom then top over i=0..ni+1, j=0..nj+1 (kmax = nk).
 * Option 2 copies from two points in, options >= 3 from one point in. */
void or_bc8(const or_cfg *c, int kind, int kmax, R *var)
{
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  const int lo = kind == 0 ? c->wbc : kind == 1 ? c->sbc : c->bbc;
  const int hi = kind == 0 ? c->ebc : kind == 1 ? c->nbc : c->tbc;
  const int sx = ni + 2, sxy = (ni + 2) * (nj + 2);
#define V8(i, j, k) var[(size_t)(i) + (size_t)sx * (size_t)(j) + (size_t)sxy * (size_t)((k) - 1)]
  if (kind == 2) {
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        if (lo == 2) V8(i, j, 1) = V8(i, j, 3);
        else if (lo >= 3) V8(i, j, 1) = V8(i, j, 2);
      }
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        if (hi == 2) V8(i, j, nk) = V8(i, j, nk - 2);
        else if (hi >= 3) V8(i, j, nk) = V8(i, j, nk - 1);
      }
    return;
  }
  for (int k = 1; k <= kmax; k++)
    for (int a = 0; a <= (kind == 0 ? nj : ni) + 1; a++) {
      if (kind == 0) {
        if (lo == 2) V8(1, a, k) = V8(3, a, k);
        else if (lo >= 3) V8(1, a, k) = V8(2, a, k);
      } else {
        if (lo == 2) V8(a, 1, k) = V8(a, 3, k);
        else if (lo >= 3) V8(a, 1, k) = V8(a, 2, k);
      }
    }
  for (int k = 1; k <= kmax; k++)
    for (int a = 0; a <= (kind == 0 ? nj : ni) + 1; a++) {
      if (kind == 0) {
        if (hi == 2) V8(ni, a, k) = V8(ni - 2, a, k);
        else if (hi >= 3) V8(ni, a, k) = V8(ni - 1, a, k);
      } else {
        if (hi == 2) V8(a, nj, k) = V8(a, nj - 2, k);
        else if (hi >= 3) V8(a, nj, k) = V8(a, nj - 1, k);
      }
    }
#undef V8
}

/* s_bcycle, bcycle.f90:159-225.  Single rank: periodic copy in x when
 * wbc=ebc=-1 over j=0..nj+1, then in y when sbc=nbc=-1 over i=0..ni+1. */
void or_bcycle(const or_cfg *c, int iwsnd, int iwrcv, int iesnd, int iercv,
               int jssnd, int jsrcv, int jnsnd, int jnrcv, int kmax, R *var)
{
  DECL_DIMS(c);
  if (c->wbc == -1 && c->ebc == -1) {
    for (int k = 1; k <= kmax; k++)
      for (int j = 0; j <= nj + 1; j++) A3(var, iwrcv, j, k) = A3(var, iesnd, j, k);
    for (int k = 1; k <= kmax; k++)
      for (int j = 0; j <= nj + 1; j++) A3(var, iercv, j, k) = A3(var, iwsnd, j, k);
  }
  if (c->sbc == -1 && c->nbc == -1) {
    for (int k = 1; k <= kmax; k++)
      for (int i = 0; i <= ni + 1; i++) A3(var, i, jsrcv, k) = A3(var, i, jnsnd, k);
    for (int k = 1; k <= kmax; k++)
      for (int i = 0; i <= ni + 1; i++) A3(var, i, jnrcv, k) = A3(var, i, jssnd, k);
  }
}

/* s_vbcs, vbcs.f90:116-134 */
void or_vbcs(const or_cfg *c, R *sf)
{
  DECL_DIMS(c);
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni - 1; i++) A3(sf, i, j, 1) = A3(sf, i, j, 2);
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni - 1; i++) A3(sf, i, j, nk - 1) = A3(sf, i, j, nk - 2);
}

/* s_vbcu, vbcu.f90:118-136 */
void or_vbcu(const or_cfg *c, R *uf)
{
  DECL_DIMS(c);
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni; i++) A3(uf, i, j, 1) = A3(uf, i, j, 2);
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni; i++) A3(uf, i, j, nk - 1) = A3(uf, i, j, nk - 2);
}

/* s_vbcv, vbcv.f90:118-136 */
void or_vbcv(const or_cfg *c, R *vf)
{
  DECL_DIMS(c);
  for (int j = 1; j <= nj; j++)
    for (int i = 1; i <= ni - 1; i++) A3(vf, i, j, 1) = A3(vf, i, j, 2);
  for (int j = 1; j <= nj; j++)
    for (int i = 1; i <= ni - 1; i++) A3(vf, i, j, nk - 1) = A3(vf, i, j, nk - 2);
}

/* s_vbcp, vbcp.f90:126-162 */
void or_vbcp(const or_cfg *c, R *ppf)
{
  DECL_DIMS(c);
  if (c->bbc == 2) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(ppf, i, j, 1) = K(2.0) * A3(ppf, i, j, 2) - A3(ppf, i, j, 3);
  } else if (c->bbc == 3) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(ppf, i, j, 1) = A3(ppf, i, j, 2);
  }
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni - 1; i++) A3(ppf, i, j, nk - 1) = A3(ppf, i, j, nk - 2);
}

/* s_vbcwc, vbcwc.f90:130-186 */
void or_vbcwc(const or_cfg *c, R *wc)
{
  DECL_DIMS(c);
  if (c->bbc == 2) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        A3(wc, i, j, 1) = -A3(wc, i, j, 3);
        A3(wc, i, j, 2) = K(0.0);
      }
  } else if (c->bbc == 3) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(wc, i, j, 1) = A3(wc, i, j, 2);
  }
  if (c->tbc == 2) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        A3(wc, i, j, nk) = -A3(wc, i, j, nk - 2);
        A3(wc, i, j, nk - 1) = K(0.0);
      }
  } else if (c->tbc >= 3) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(wc, i, j, nk) = A3(wc, i, j, nk - 1);
  }
}

/* s_vbcw, vbcw.f90:201-499 for flat terrain (j31=j32=0, mfcopt=0): the
 * terrain-slope terms reduce to exact zeros, so w(1) = -jcb8w(3)*wc(3) - 0,
 * w(2) = -.25*0 = -0, w(nk) = -jcb8w(nk-2)*wc(nk-2) - 0, w(nk-1) = 0. */
void or_vbcw(const or_cfg *c, const R *jcb8w, const R *wc, R *wf)
{
  DECL_DIMS(c);
  const R zero = K(0.0);
  if (c->bbc == 2) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        R t = K(0.25) * ((zero + zero) + (zero + zero));
        A3(wf, i, j, 1) = -A3(jcb8w, i, j, 3) * A3(wc, i, j, 3) - t;
      }
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(wf, i, j, 2) = -K(0.25) * ((zero + zero) + (zero + zero));
  } else if (c->bbc == 3) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(wf, i, j, 1) = A3(wf, i, j, 2);
  }
  if (c->tbc == 2) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        R t = K(0.25) * ((zero + zero) + (zero + zero));
        A3(wf, i, j, nk) = -A3(jcb8w, i, j, nk - 2) * A3(wc, i, j, nk - 2) - t;
      }
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(wf, i, j, nk - 1) = K(0.0);
  } else if (c->tbc == 3) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(wf, i, j, nk) = A3(wf, i, j, nk - 1);
  }
}

#include "cress_obc.c"
#include "cress_smooth.c"

/* s_phy2cnt, phy2cnt.f90:193-227 (trnopt=0) then vbcwc (:411).  The
 * advopt>=4 halo branch (:350-405) is not taken for advopt<=3. */
void or_phy2cnt(const or_cfg *c, int sthopt, const R *jcb8w, const R *w, R *wc)
{
  DECL_DIMS(c);
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(wc, i, j, k) = sthopt == 0 ? A3(w, i, j, k) : A3(w, i, j, k) / A3(jcb8w, i, j, k);
  /* the lateral fills of wc (:396-405) run only for advopt>=4 */
  or_vbcwc(c, wc);
}

/* ------------------------------------------------------------------------ */
/* Large-step forcing                                                       */
/* ------------------------------------------------------------------------ */

/* s_rstuvwc, rstuvwc.f90:172-210 (mfcopt=0) */
void or_rstuvwc(const or_cfg *c, const R *rst8u, const R *rst8v, const R *rst8w,
                const R *u, const R *v, const R *wc, R *rstxu, R *rstxv, R *rstxwc)
{
  DECL_DIMS(c);
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = c->iwest; i <= ni + 1 - c->ieast; i++)
        A3(rstxu, i, j, k) = A3(rst8u, i, j, k) * A3(u, i, j, k);
    for (int j = c->jsouth; j <= nj + 1 - c->jnorth; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(rstxv, i, j, k) = A3(rst8v, i, j, k) * A3(v, i, j, k);
  }
  OMP_FOR
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(rstxwc, i, j, k) = A3(rst8w, i, j, k) * A3(wc, i, j, k);
}

/* s_advs, advs.f90:205-445.  sfrc is overwritten; vadv, t1..t3 scratch. */
void or_advs(const or_cfg *c, const R *rstxu, const R *rstxv, const R *rstxwc,
             const R *s, R *sfrc, R *vadv, R *tmp1, R *tmp2, R *tmp3)
{
  DECL_DIMS(c);
  const int advopt = c->advopt;
  const int iw = c->iwest, ie = c->ieast, js = c->jsouth, jn = c->jnorth;
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy, dziv = (R)1 / (R)c->dz;
  const R dxv05n = K(-0.5) * dxiv, dyv05n = K(-0.5) * dyiv, dzv05n = K(-0.5) * dziv;
  const R dxv24 = ONED24 * dxiv, dyv24 = ONED24 * dyiv, dzv24 = ONED24 * dziv;
  if (advopt > 3) {
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++)
      for (int j = 2; j <= nj - 2; j++)
        for (int i = 2; i <= ni - 2; i++) A3(sfrc, i, j, k) = K(0.0);
    return;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(tmp1, i, j, k) = A3(rstxu, i, j, k) * (A3(s, i, j, k) - A3(s, i - 1, j, k)) * dxv05n;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp2, i, j, k) = A3(rstxv, i, j, k) * (A3(s, i, j, k) - A3(s, i, j - 1, k)) * dyv05n;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp3, i, j, k) = A3(rstxwc, i, j, k) * (A3(s, i, j, k) - A3(s, i, j, k - 1)) * dzv05n;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (advopt == 1) {
          A3(sfrc, i, j, k) = (A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1)) +
                              ((A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k)));
        } else {
          A3(sfrc, i, j, k) = (A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                              (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k));
          if (advopt == 3) A3(vadv, i, j, k) = A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1);
        }
      }
  if (advopt == 1) return;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 1 + iw; i <= ni - 1 - ie; i++)
        A3(tmp1, i, j, k) = (A3(rstxu, i, j, k) + A3(rstxu, i + 1, j, k)) *
                            (A3(s, i + 1, j, k) - A3(s, i - 1, j, k)) * dxv24;
    for (int j = 1 + js; j <= nj - 1 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(tmp2, i, j, k) = (A3(rstxv, i, j, k) + A3(rstxv, i, j + 1, k)) *
                            (A3(s, i, j + 1, k) - A3(s, i, j - 1, k)) * dyv24;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(sfrc, i, j, k) = FOURD3 * A3(sfrc, i, j, k) +
                            ((A3(tmp1, i - 1, j, k) + A3(tmp1, i + 1, j, k)) +
                             (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j + 1, k)));
  if (advopt == 2) {
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++)
      for (int j = 2; j <= nj - 2; j++)
        for (int i = 2; i <= ni - 2; i++)
          A3(sfrc, i, j, k) = A3(sfrc, i, j, k) + (A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1));
    return;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(tmp3, i, j, k) = (A3(rstxwc, i, j, k) + A3(rstxwc, i, j, k + 1)) *
                            (A3(s, i, j, k + 1) - A3(s, i, j, k - 1)) * dzv24;
  OMP_FOR
  for (int k = 3; k <= nk - 3; k++)
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(vadv, i, j, k) = FOURD3 * A3(vadv, i, j, k) + (A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k + 1));
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) A3(sfrc, i, j, k) = A3(sfrc, i, j, k) + A3(vadv, i, j, k);
}

/* s_advuvw, advuvw.f90:209-902.  ufrc/vfrc/wfrc accumulate (+=). */
void or_advuvw(const or_cfg *c, const R *rstxu, const R *rstxv, const R *rstxwc,
               const R *u, const R *v, const R *w, R *ufrc, R *vfrc, R *wfrc,
               R *hadv, R *vadv, R *tmp1, R *tmp2, R *tmp3)
{
  DECL_DIMS(c);
  const int advopt = c->advopt;
  const int iw = c->iwest, ie = c->ieast, js = c->jsouth, jn = c->jnorth;
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy, dziv = (R)1 / (R)c->dz;
  const R dxv25n = K(-0.25) * dxiv, dyv25n = K(-0.25) * dyiv, dzv25n = K(-0.25) * dziv;
  const R dxv24 = ONED24 * dxiv, dyv24 = ONED24 * dyiv, dzv24 = ONED24 * dziv;
  const R dxv48 = K(0.5) * dxv24, dyv48 = K(0.5) * dyv24, dzv48 = K(0.5) * dzv24;
  (void)dxv48;

  /* ---- u (advuvw.f90:231-448) ---- */
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(tmp1, i, j, k) = (A3(rstxu, i, j, k) + A3(rstxu, i + 1, j, k)) *
                            (A3(u, i + 1, j, k) - A3(u, i, j, k)) * dxv25n;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(tmp2, i, j, k) = (A3(rstxv, i - 1, j, k) + A3(rstxv, i, j, k)) *
                            (A3(u, i, j, k) - A3(u, i, j - 1, k)) * dyv25n;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(tmp3, i, j, k) = (A3(rstxwc, i - 1, j, k) + A3(rstxwc, i, j, k)) *
                            (A3(u, i, j, k) - A3(u, i, j, k - 1)) * dzv25n;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++) {
        if (advopt == 1) {
          A3(ufrc, i, j, k) = A3(ufrc, i, j, k) + ((A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1)) +
                              ((A3(tmp1, i - 1, j, k) + A3(tmp1, i, j, k)) +
                               (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k))));
        } else {
          A3(hadv, i, j, k) = (A3(tmp1, i - 1, j, k) + A3(tmp1, i, j, k)) +
                              (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k));
          if (advopt == 3) A3(vadv, i, j, k) = A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1);
        }
      }
  if (advopt >= 2) {
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++) {
      for (int j = 2 + js; j <= nj - 2 - jn; j++)
        for (int i = 1 + iw; i <= ni - ie; i++)
          A3(tmp1, i, j, k) = (A3(rstxu, i - 1, j, k) + A3(rstxu, i + 1, j, k)) *
                              (A3(u, i + 1, j, k) - A3(u, i - 1, j, k)) * dxv24;
      for (int j = 1 + js; j <= nj - 1 - jn; j++)
        for (int i = 2 + iw; i <= ni - 1 - ie; i++)
          A3(tmp2, i, j, k) = ((A3(rstxv, i - 1, j, k) + A3(rstxv, i, j, k)) +
                               (A3(rstxv, i - 1, j + 1, k) + A3(rstxv, i, j + 1, k))) *
                              (A3(u, i, j + 1, k) - A3(u, i, j - 1, k)) * dyv48;
    }
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++)
      for (int j = 2 + js; j <= nj - 2 - jn; j++)
        for (int i = 2 + iw; i <= ni - 1 - ie; i++)
          A3(hadv, i, j, k) = FOURD3 * A3(hadv, i, j, k) +
                              ((A3(tmp1, i - 1, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j + 1, k)));
    if (advopt == 2) {
      for (int k = 2; k <= nk - 2; k++)
        for (int j = 2; j <= nj - 2; j++)
          for (int i = 2; i <= ni - 1; i++)
            A3(ufrc, i, j, k) = A3(ufrc, i, j, k) + A3(hadv, i, j, k) +
                                (A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1));
    } else {
      for (int k = 2; k <= nk - 2; k++)
        for (int j = 2 + js; j <= nj - 2 - jn; j++)
          for (int i = 2 + iw; i <= ni - 1 - ie; i++)
            A3(tmp3, i, j, k) = ((A3(rstxwc, i - 1, j, k) + A3(rstxwc, i, j, k)) +
                                 (A3(rstxwc, i - 1, j, k + 1) + A3(rstxwc, i, j, k + 1))) *
                                (A3(u, i, j, k + 1) - A3(u, i, j, k - 1)) * dzv48;
      for (int k = 3; k <= nk - 3; k++)
        for (int j = 2 + js; j <= nj - 2 - jn; j++)
          for (int i = 2 + iw; i <= ni - 1 - ie; i++)
            A3(vadv, i, j, k) = FOURD3 * A3(vadv, i, j, k) + (A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k + 1));
      for (int k = 2; k <= nk - 2; k++)
        for (int j = 2; j <= nj - 2; j++)
          for (int i = 2; i <= ni - 1; i++)
            A3(ufrc, i, j, k) = A3(ufrc, i, j, k) + A3(hadv, i, j, k) + A3(vadv, i, j, k);
    }
  }

  /* ---- v (advuvw.f90:458-675) ---- */
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(tmp1, i, j, k) = (A3(rstxu, i, j - 1, k) + A3(rstxu, i, j, k)) *
                            (A3(v, i, j, k) - A3(v, i - 1, j, k)) * dxv25n;
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp2, i, j, k) = (A3(rstxv, i, j, k) + A3(rstxv, i, j + 1, k)) *
                            (A3(v, i, j + 1, k) - A3(v, i, j, k)) * dyv25n;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp3, i, j, k) = (A3(rstxwc, i, j - 1, k) + A3(rstxwc, i, j, k)) *
                            (A3(v, i, j, k) - A3(v, i, j, k - 1)) * dzv25n;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (advopt == 1) {
          A3(vfrc, i, j, k) = A3(vfrc, i, j, k) + ((A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1)) +
                              ((A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j, k))));
        } else {
          A3(hadv, i, j, k) = (A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                              (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j, k));
          if (advopt == 3) A3(vadv, i, j, k) = A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1);
        }
      }
  if (advopt >= 2) {
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++) {
      for (int j = 2 + js; j <= nj - 1 - jn; j++)
        for (int i = 1 + iw; i <= ni - 1 - ie; i++)
          A3(tmp1, i, j, k) = ((A3(rstxu, i, j - 1, k) + A3(rstxu, i, j, k)) +
                               (A3(rstxu, i + 1, j - 1, k) + A3(rstxu, i + 1, j, k))) *
                              (A3(v, i + 1, j, k) - A3(v, i - 1, j, k)) * dxv48;
      for (int j = 1 + js; j <= nj - jn; j++)
        for (int i = 2 + iw; i <= ni - 2 - ie; i++)
          A3(tmp2, i, j, k) = (A3(rstxv, i, j - 1, k) + A3(rstxv, i, j + 1, k)) *
                              (A3(v, i, j + 1, k) - A3(v, i, j - 1, k)) * dyv24;
    }
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++)
      for (int j = 2 + js; j <= nj - 1 - jn; j++)
        for (int i = 2 + iw; i <= ni - 2 - ie; i++)
          A3(hadv, i, j, k) = FOURD3 * A3(hadv, i, j, k) +
                              ((A3(tmp1, i - 1, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j + 1, k)));
    if (advopt == 2) {
      for (int k = 2; k <= nk - 2; k++)
        for (int j = 2; j <= nj - 1; j++)
          for (int i = 2; i <= ni - 2; i++)
            A3(vfrc, i, j, k) = A3(vfrc, i, j, k) + A3(hadv, i, j, k) +
                                (A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1));
    } else {
      for (int k = 2; k <= nk - 2; k++)
        for (int j = 2 + js; j <= nj - 1 - jn; j++)
          for (int i = 2 + iw; i <= ni - 2 - ie; i++)
            A3(tmp3, i, j, k) = ((A3(rstxwc, i, j - 1, k) + A3(rstxwc, i, j, k)) +
                                 (A3(rstxwc, i, j - 1, k + 1) + A3(rstxwc, i, j, k + 1))) *
                                (A3(v, i, j, k + 1) - A3(v, i, j, k - 1)) * dzv48;
      for (int k = 3; k <= nk - 3; k++)
        for (int j = 2 + js; j <= nj - 1 - jn; j++)
          for (int i = 2 + iw; i <= ni - 2 - ie; i++)
            A3(vadv, i, j, k) = FOURD3 * A3(vadv, i, j, k) + (A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k + 1));
      for (int k = 2; k <= nk - 2; k++)
        for (int j = 2; j <= nj - 1; j++)
          for (int i = 2; i <= ni - 2; i++)
            A3(vfrc, i, j, k) = A3(vfrc, i, j, k) + A3(hadv, i, j, k) + A3(vadv, i, j, k);
    }
  }

  /* ---- w (advuvw.f90:685-902) ---- */
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(tmp1, i, j, k) = (A3(rstxu, i, j, k - 1) + A3(rstxu, i, j, k)) *
                            (A3(w, i, j, k) - A3(w, i - 1, j, k)) * dxv25n;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp2, i, j, k) = (A3(rstxv, i, j, k - 1) + A3(rstxv, i, j, k)) *
                            (A3(w, i, j, k) - A3(w, i, j - 1, k)) * dyv25n;
  }
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp3, i, j, k) = (A3(rstxwc, i, j, k) + A3(rstxwc, i, j, k + 1)) *
                            (A3(w, i, j, k + 1) - A3(w, i, j, k)) * dzv25n;
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (advopt == 1) {
          A3(wfrc, i, j, k) = A3(wfrc, i, j, k) + ((A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k)) +
                              ((A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k))));
        } else {
          A3(hadv, i, j, k) = (A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                              (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k));
          if (advopt == 3) A3(vadv, i, j, k) = A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k);
        }
      }
  if (advopt >= 2) {
    OMP_FOR
    for (int k = 2; k <= nk - 1; k++) {
      for (int j = 2 + js; j <= nj - 2 - jn; j++)
        for (int i = 1 + iw; i <= ni - 1 - ie; i++)
          A3(tmp1, i, j, k) = ((A3(rstxu, i, j, k - 1) + A3(rstxu, i, j, k)) +
                               (A3(rstxu, i + 1, j, k - 1) + A3(rstxu, i + 1, j, k))) *
                              (A3(w, i + 1, j, k) - A3(w, i - 1, j, k)) * dxv48;
      for (int j = 1 + js; j <= nj - 1 - jn; j++)
        for (int i = 2 + iw; i <= ni - 2 - ie; i++)
          A3(tmp2, i, j, k) = ((A3(rstxv, i, j, k - 1) + A3(rstxv, i, j, k)) +
                               (A3(rstxv, i, j + 1, k - 1) + A3(rstxv, i, j + 1, k))) *
                              (A3(w, i, j + 1, k) - A3(w, i, j - 1, k)) * dyv48;
    }
    OMP_FOR
    for (int k = 2; k <= nk - 1; k++)
      for (int j = 2 + js; j <= nj - 2 - jn; j++)
        for (int i = 2 + iw; i <= ni - 2 - ie; i++)
          A3(hadv, i, j, k) = FOURD3 * A3(hadv, i, j, k) +
                              ((A3(tmp1, i - 1, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j + 1, k)));
    if (advopt == 2) {
      for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 2; j++)
          for (int i = 2; i <= ni - 2; i++)
            A3(wfrc, i, j, k) = A3(wfrc, i, j, k) + A3(hadv, i, j, k) +
                                (A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k));
    } else {
      for (int k = 2; k <= nk - 1; k++)
        for (int j = 2 + js; j <= nj - 2 - jn; j++)
          for (int i = 2 + iw; i <= ni - 2 - ie; i++)
            A3(tmp3, i, j, k) = (A3(rstxwc, i, j, k - 1) + A3(rstxwc, i, j, k + 1)) *
                                (A3(w, i, j, k + 1) - A3(w, i, j, k - 1)) * dzv24;
      for (int k = 3; k <= nk - 2; k++)
        for (int j = 2 + js; j <= nj - 2 - jn; j++)
          for (int i = 2 + iw; i <= ni - 2 - ie; i++)
            A3(vadv, i, j, k) = FOURD3 * A3(vadv, i, j, k) + (A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k + 1));
      for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 2; j++)
          for (int i = 2; i <= ni - 2; i++)
            A3(wfrc, i, j, k) = A3(wfrc, i, j, k) + A3(hadv, i, j, k) + A3(vadv, i, j, k);
    }
  }
}

/* s_advp, advp.f90:249-785 (mfcopt=0, diaopt=0).  pfrc is overwritten. */
void or_advp(const or_cfg *c, const R *jcb8u, const R *jcb8v, const R *jcb8w,
             const R *u, const R *v, const R *wc, const R *pp, R *pfrc,
             R *jcbxu, R *jcbxv, R *jcbxwc, R *vadv, R *tmp1, R *tmp2, R *tmp3)
{
  DECL_DIMS(c);
  const int advopt = c->advopt;
  const int iw = c->iwest, ie = c->ieast, js = c->jsouth, jn = c->jnorth;
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy, dziv = (R)1 / (R)c->dz;
  const R dxv05n = K(-0.5) * dxiv, dyv05n = K(-0.5) * dyiv, dzv05n = K(-0.5) * dziv;
  const R dxv24 = ONED24 * dxiv, dyv24 = ONED24 * dyiv, dzv24 = ONED24 * dziv;
  if (advopt > 3) {
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++)
      for (int j = 2; j <= nj - 2; j++)
        for (int i = 2; i <= ni - 2; i++) A3(pfrc, i, j, k) = K(0.0);
    return;
  }
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni; i++) A3(jcbxu, i, j, k) = A3(jcb8u, i, j, k) * A3(u, i, j, k);
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni - 1; i++) A3(jcbxv, i, j, k) = A3(jcb8v, i, j, k) * A3(v, i, j, k);
  }
  OMP_FOR
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(jcbxwc, i, j, k) = A3(jcb8w, i, j, k) * A3(wc, i, j, k);
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(tmp1, i, j, k) = A3(jcbxu, i, j, k) * (A3(pp, i, j, k) - A3(pp, i - 1, j, k)) * dxv05n;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp2, i, j, k) = A3(jcbxv, i, j, k) * (A3(pp, i, j, k) - A3(pp, i, j - 1, k)) * dyv05n;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp3, i, j, k) = A3(jcbxwc, i, j, k) * (A3(pp, i, j, k) - A3(pp, i, j, k - 1)) * dzv05n;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (advopt == 1) {
          A3(pfrc, i, j, k) = (A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1)) +
                              ((A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                               (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k)));
        } else {
          A3(pfrc, i, j, k) = (A3(tmp1, i, j, k) + A3(tmp1, i + 1, j, k)) +
                              (A3(tmp2, i, j, k) + A3(tmp2, i, j + 1, k));
          if (advopt == 3) A3(vadv, i, j, k) = A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1);
        }
      }
  if (advopt == 1) return;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 1 + iw; i <= ni - 1 - ie; i++)
        A3(tmp1, i, j, k) = (A3(jcbxu, i, j, k) + A3(jcbxu, i + 1, j, k)) *
                            (A3(pp, i + 1, j, k) - A3(pp, i - 1, j, k)) * dxv24;
    for (int j = 1 + js; j <= nj - 1 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(tmp2, i, j, k) = (A3(jcbxv, i, j, k) + A3(jcbxv, i, j + 1, k)) *
                            (A3(pp, i, j + 1, k) - A3(pp, i, j - 1, k)) * dyv24;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(pfrc, i, j, k) = FOURD3 * A3(pfrc, i, j, k) +
                            ((A3(tmp1, i - 1, j, k) + A3(tmp1, i + 1, j, k)) +
                             (A3(tmp2, i, j - 1, k) + A3(tmp2, i, j + 1, k)));
  if (advopt == 2) {
    OMP_FOR
    for (int k = 2; k <= nk - 2; k++)
      for (int j = 2; j <= nj - 2; j++)
        for (int i = 2; i <= ni - 2; i++)
          A3(pfrc, i, j, k) = A3(pfrc, i, j, k) + (A3(tmp3, i, j, k) + A3(tmp3, i, j, k + 1));
    return;
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(tmp3, i, j, k) = (A3(jcbxwc, i, j, k) + A3(jcbxwc, i, j, k + 1)) *
                            (A3(pp, i, j, k + 1) - A3(pp, i, j, k - 1)) * dzv24;
  OMP_FOR
  for (int k = 3; k <= nk - 3; k++)
    for (int j = 2 + js; j <= nj - 2 - jn; j++)
      for (int i = 2 + iw; i <= ni - 2 - ie; i++)
        A3(vadv, i, j, k) = FOURD3 * A3(vadv, i, j, k) + (A3(tmp3, i, j, k - 1) + A3(tmp3, i, j, k + 1));
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) A3(pfrc, i, j, k) = A3(pfrc, i, j, k) + A3(vadv, i, j, k);
}

/* s_coriuv, coriuv.f90:124-182 */
void or_coriuv(const or_cfg *c, const R *fc, const R *rst, const R *u, const R *v,
               R *ufrc, R *vfrc, R *tmp1)
{
  DECL_DIMS(c);
  /* fc(0:ni+1,0:nj+1,1:2): component 1 is read */
#define FC1(i, j) fc[(size_t)(i) + NX * (size_t)(j)]
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(tmp1, i, j, k) = FC1(i, j) * A3(rst, i, j, k) * (A3(v, i, j, k) + A3(v, i, j + 1, k));
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(ufrc, i, j, k) = A3(ufrc, i, j, k) + (A3(tmp1, i - 1, j, k) + A3(tmp1, i, j, k));
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tmp1, i, j, k) = -FC1(i, j) * A3(rst, i, j, k) * (A3(u, i, j, k) + A3(u, i + 1, j, k));
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(vfrc, i, j, k) = A3(vfrc, i, j, k) + (A3(tmp1, i, j - 1, k) + A3(tmp1, i, j, k));
#undef FC1
}

/* s_buoywb, buoywb.f90:152-319, moist, gwmopt=0, cphopt!=0. */
void or_buoywb(const or_cfg *c, const R *ptbr, const R *qvbr, const R *rst,
               const R *ptp, const R *qv, const R *qall, R *wfrc, R *qvd, R *wb8s)
{
  DECL_DIMS(c);
  const R g05 = K(0.5) * GRAV, epsva = K(0.622);
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) A3(qvd, i, j, k) = A3(qv, i, j, k) - A3(qvbr, i, j, k);
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (c->cphopt == 0)
          A3(wb8s, i, j, k) = g05 * A3(rst, i, j, k) *
              (A3(qvd, i, j, k) / (epsva + A3(qvbr, i, j, k)) -
               A3(qvd, i, j, k) / (K(1.0) + A3(qvbr, i, j, k)) + A3(ptp, i, j, k) / A3(ptbr, i, j, k));
        else
          A3(wb8s, i, j, k) = g05 * A3(rst, i, j, k) *
              (A3(qvd, i, j, k) / (epsva + A3(qvbr, i, j, k)) -
               (A3(qvd, i, j, k) + A3(qall, i, j, k)) / (K(1.0) + A3(qvbr, i, j, k)) +
               A3(ptp, i, j, k) / A3(ptbr, i, j, k));
      }
  OMP_FOR
  for (int k = 3; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(wfrc, i, j, k) = A3(wfrc, i, j, k) + (A3(wb8s, i, j, k - 1) + A3(wb8s, i, j, k));
}

/* rbct(1:ni,1:nj,1:nk,1:2): the sponge profile (vspdmp.f90:269-285) */
#define RBCT2(i, j, k) rbct[((size_t)(i) - 1) + (size_t)ni * (((size_t)(j) - 1) + (size_t)nj * (((size_t)(k) - 1) + (size_t)nk))]

/* s_vspuvw, vspuvw.f90:212-408, vspopt=2 branch (relaxation to the base
 * state, rbct(:,:,:,2)).  rbct8s is scratch. */
void or_vspuvw(const or_cfg *c, const R *ubr, const R *vbr, const R *rst8u,
               const R *rst8v, const R *rst8w, const R *up, const R *vp, const R *wp,
               const R *rbct, R *ufrc, R *vfrc, R *wfrc, R *rbct8s)
{
  DECL_DIMS(c);
  const int k0 = c->ksp0[1];
  const char *vv = c->vspvar;
  if (c->vspopt != 2) return;
  if (vv[0] == 'o' || vv[1] == 'o') {
    OMP_FOR
    for (int k = k0 - 1; k <= nk - 2; k++)
      for (int j = 1; j <= nj - 1; j++)
        for (int i = 1; i <= ni - 1; i++)
          A3(rbct8s, i, j, k) = K(0.25) * (RBCT2(i, j, k) + RBCT2(i, j, k + 1));
  }
  if (vv[0] == 'o') {
    OMP_FOR
    for (int k = k0 - 1; k <= nk - 2; k++)
      for (int j = 2; j <= nj - 2; j++)
        for (int i = 2; i <= ni - 1; i++)
          A3(ufrc, i, j, k) = A3(ufrc, i, j, k) - A3(rst8u, i, j, k) *
              (A3(rbct8s, i - 1, j, k) + A3(rbct8s, i, j, k)) * (A3(up, i, j, k) - A3(ubr, i, j, k));
  }
  if (vv[1] == 'o') {
    OMP_FOR
    for (int k = k0 - 1; k <= nk - 2; k++)
      for (int j = 2; j <= nj - 1; j++)
        for (int i = 2; i <= ni - 2; i++)
          A3(vfrc, i, j, k) = A3(vfrc, i, j, k) - A3(rst8v, i, j, k) *
              (A3(rbct8s, i, j - 1, k) + A3(rbct8s, i, j, k)) * (A3(vp, i, j, k) - A3(vbr, i, j, k));
  }
  if (vv[2] == 'o') {
    OMP_FOR
    for (int k = k0; k <= nk - 1; k++)
      for (int j = 2; j <= nj - 2; j++)
        for (int i = 2; i <= ni - 2; i++)
          A3(wfrc, i, j, k) = A3(wfrc, i, j, k) - RBCT2(i, j, k) * A3(rst8w, i, j, k) * A3(wp, i, j, k);
  }
}

/* s_vsps (vspopt=2, vsps.f90:183-197) / s_vsps0 (vsps0.f90:114-127):
 * sfrc -= .5*(rbct_k+rbct_k+1)*rst*sp.  s_vspqv (vspqv.f90:185-200) and
 * s_vspp (vspp.f90:180-194) relax to qvbr and use jcb; ref may be NULL. */
void or_vsps(const or_cfg *c, const R *rst, const R *sp, const R *ref, const R *rbct, R *sfrc)
{
  DECL_DIMS(c);
  OMP_FOR
  for (int k = c->ksp0[1] - 1; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        R d = ref ? (A3(sp, i, j, k) - A3(ref, i, j, k)) : A3(sp, i, j, k);
        if (ref)
          A3(sfrc, i, j, k) = A3(sfrc, i, j, k) - K(0.5) * (RBCT2(i, j, k) + RBCT2(i, j, k + 1)) *
                              A3(rst, i, j, k) * d;
        else
          A3(sfrc, i, j, k) = A3(sfrc, i, j, k) - K(0.5) * (RBCT2(i, j, k) + RBCT2(i, j, k + 1)) *
                              A3(rst, i, j, k) * d;
      }
}

/* s_advbspt, advbspt.f90:143-190 (gwmopt=0: accumulate). */
void or_advbspt(const or_cfg *c, const R *ptbr, const R *rbr, const R *w, R *ptadv, R *pta8w)
{
  DECL_DIMS(c);
  const R dziv = (R)1 / (R)c->dz;
  const R dziv25 = K(0.25) * dziv;
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(pta8w, i, j, k) = (A3(rbr, i, j, k - 1) + A3(rbr, i, j, k)) * A3(w, i, j, k) *
                             (A3(ptbr, i, j, k - 1) - A3(ptbr, i, j, k)) * dziv25;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(ptadv, i, j, k) = A3(ptadv, i, j, k) + (A3(pta8w, i, j, k) + A3(pta8w, i, j, k + 1));
}

/* s_totalqwi, totalqwi.f90:153-170 (cphopt=2, haiopt=0).  qw = (qc,qr) and
 * qi = (qi,qs,qg) are species-stacked 3-D blocks. */
void or_totalqwi(const or_cfg *c, const R *qw, const R *qi, R *qall)
{
  DECL_DIMS(c);
  const size_t n = (size_t)nfield(c);
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        size_t p = (size_t)i + NX * ((size_t)j + NY * ((size_t)k - 1));
        qall[p] = qw[p] + qw[p + n] + qi[p] + qi[p + n] + qi[p + 2 * n];
      }
}

/* ------------------------------------------------------------------------ */
/* Acoustic substep                                                          */
/* ------------------------------------------------------------------------ */

/* s_diver3d, diver3d.f90:182-326 (mfcopt=0).  div3d is the output; s1..s3
 * are the flux scratch arrays (the reference's tmp1..tmp3). */
void or_diver3d(const or_cfg *c, const R *rst8u, const R *rst8v, const R *rst8w,
                const R *u, const R *v, const R *wc, R *div3d, R *s1, R *s2, R *s3)
{
  DECL_DIMS(c);
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy, dziv = (R)1 / (R)c->dz;
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni; i++) A3(s1, i, j, k) = A3(rst8u, i, j, k) * A3(u, i, j, k);
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni - 1; i++) A3(s2, i, j, k) = A3(rst8v, i, j, k) * A3(v, i, j, k);
  }
  OMP_FOR
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(s3, i, j, k) = A3(rst8w, i, j, k) * A3(wc, i, j, k);
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(div3d, i, j, k) = (A3(s3, i, j, k) - A3(s3, i, j, k + 1)) * dziv +
                             ((A3(s1, i, j, k) - A3(s1, i + 1, j, k)) * dxiv +
                              (A3(s2, i, j, k) - A3(s2, i, j + 1, k)) * dyiv);
}

/* s_pgrad, pgrad.f90:251-395 (trnopt=0, mfcopt=0, divopt 0/1).  The
 * divergence lands in tmp1 (diver3d's div3d, :262-264), tmp3 holds the
 * damped pressure p + divcv*jcb*div (:294-306), then jcb*p (:357-369). */
void or_pgrad(const or_cfg *c, R dts, const R *jcb, const R *rst8u, const R *rst8v,
              const R *rst8w, const R *u, const R *v, const R *wc, const R *pp,
              R *upg, R *vpg, R *wpg, R *tmp1, R *tmp2, R *tmp3)
{
  DECL_DIMS(c);
  const R dz = (R)c->dz;
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy, dziv = (R)1 / (R)c->dz;
  const R divndc = K(0.05);
  const R divcv = divndc / dts * dz * dz;
  if (c->divopt >= 1)
    or_diver3d(c, rst8u, rst8v, rst8w, u, v, wc, tmp1, tmp2, tmp3, wpg);
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(tmp3, i, j, k) = c->divopt == 0 ? A3(pp, i, j, k)
                            : A3(pp, i, j, k) + divcv * A3(jcb, i, j, k) * A3(tmp1, i, j, k);
  OMP_FOR
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(wpg, i, j, k) = (A3(tmp3, i, j, k - 1) - A3(tmp3, i, j, k)) * dziv;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(tmp3, i, j, k) = A3(jcb, i, j, k) * A3(tmp3, i, j, k);
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(upg, i, j, k) = (A3(tmp3, i - 1, j, k) - A3(tmp3, i, j, k)) * dxiv;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(vpg, i, j, k) = (A3(tmp3, i, j - 1, k) - A3(tmp3, i, j, k)) * dyiv;
  }
}

/* s_diverpih, diverpih.f90:188-214 (trnopt=0) with s_diver2d,
 * diver2d.f90:166-295 (mfcopt=0): pdiv = rcsq * div_h(jcb8 . u). */
void or_diverpih(const or_cfg *c, const R *jcb8u, const R *jcb8v, const R *rcsq,
                 const R *u, const R *v, R *pdiv, R *tmp1, R *tmp2)
{
  DECL_DIMS(c);
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy;
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni; i++) A3(tmp1, i, j, k) = A3(jcb8u, i, j, k) * A3(u, i, j, k);
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni - 1; i++) A3(tmp2, i, j, k) = A3(jcb8v, i, j, k) * A3(v, i, j, k);
  }
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++)
        A3(pdiv, i, j, k) = (A3(tmp1, i, j, k) - A3(tmp1, i + 1, j, k)) * dxiv +
                            (A3(tmp2, i, j, k) - A3(tmp2, i, j + 1, k)) * dyiv;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) A3(pdiv, i, j, k) = A3(rcsq, i, j, k) * A3(pdiv, i, j, k);
}

/* s_diverpiv, diverpiv.f90:118-130 */
void or_diverpiv(const or_cfg *c, const R *rcsq, const R *w, R *pdiv)
{
  DECL_DIMS(c);
  const R dziv = (R)1 / (R)c->dz;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(pdiv, i, j, k) = A3(rcsq, i, j, k) * (A3(w, i, j, k) - A3(w, i, j, k + 1)) * dziv;
}

/* s_advbspi, advbspi.f90:140-184.  fore=0: 'back', fore=1: 'fore'. */
void or_advbspi(const or_cfg *c, int fore, const R *rst, const R *w, const R *pfrc,
                const R *phdiv, const R *pvdiv, R *fp)
{
  DECL_DIMS(c);
  const R weicoe = (R)c->weicoe;
  const R weic1m = K(1.0) - weicoe;
  const R g05 = K(0.5) * GRAV;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (!fore)
          A3(fp, i, j, k) = A3(pfrc, i, j, k) + A3(phdiv, i, j, k) +
              weic1m * (g05 * (A3(w, i, j, k) + A3(w, i, j, k + 1)) * A3(rst, i, j, k) + A3(pvdiv, i, j, k));
        else
          A3(fp, i, j, k) = A3(fp, i, j, k) +
              weicoe * (g05 * (A3(w, i, j, k) + A3(w, i, j, k + 1)) * A3(rst, i, j, k) + A3(pvdiv, i, j, k));
      }
}

/* s_pgradiv, pgradiv.f90:135-170 */
void or_pgradiv(const or_cfg *c, R dts, const R *jcb, const R *wfrc, const R *fp,
                R *fw, R *fpdvj)
{
  DECL_DIMS(c);
  const R dziv = (R)1 / (R)c->dz;
  const R dtdzw = dts * dziv * (R)c->weicoe;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) A3(fpdvj, i, j, k) = A3(fp, i, j, k) / A3(jcb, i, j, k);
  OMP_FOR
  for (int k = 3; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(fw, i, j, k) = A3(wfrc, i, j, k) + A3(fw, i, j, k) +
                          (A3(fpdvj, i, j, k - 1) - A3(fpdvj, i, j, k)) * dtdzw;
}

/* s_buoywsi, buoywsi.f90:153-211 (gwmopt=0) */
void or_buoywsi(const or_cfg *c, R dts, const R *rbr, const R *rst, const R *rcsq,
                const R *pp, const R *fp, R *fw, R *wb8s)
{
  DECL_DIMS(c);
  const R dtw = dts * (R)c->weicoe;
  const R g05n = K(-0.5) * GRAV;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(wb8s, i, j, k) = (A3(pp, i, j, k) * A3(rst, i, j, k) + A3(fp, i, j, k) * A3(rbr, i, j, k) * dtw) /
                            A3(rcsq, i, j, k) * g05n;
  OMP_FOR
  for (int k = 3; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(fw, i, j, k) = A3(fw, i, j, k) + (A3(wb8s, i, j, k - 1) + A3(wb8s, i, j, k));
}

/* s_gaussel, gaussel.f90:169-262 (impopt=1): Thomas algorithm over k for each
 * column (i,j); rr/ss/tt are sub/diagonal/super, ff the right-hand side and
 * solution.  ss is overwritten with the normalised super-diagonal. */
void or_gaussel(const or_cfg *c, int istr, int iend, int jstr, int jend, int kstr,
                int kend, R *rr, R *ss, R *tt, R *ff)
{
  DECL_DIMS(c);
  const int kem1 = kend - 1;
  if (kstr == kend) {
    for (int j = jstr; j <= jend; j++)
      for (int i = istr; i <= iend; i++) A3(ff, i, j, kstr) = A3(ff, i, j, kstr) / A3(ss, i, j, kstr);
  } else {
    for (int j = jstr; j <= jend; j++)
      for (int i = istr; i <= iend; i++) {
        R aa = K(1.0) / A3(ss, i, j, kstr);
        A3(ss, i, j, kstr) = A3(tt, i, j, kstr) * aa;
        A3(ff, i, j, kstr) = A3(ff, i, j, kstr) * aa;
      }
  }
  /* the recurrences run in k; columns (j) are independent */
  if (kstr + 2 <= kend) {
    OMP_FOR
    for (int j = jstr; j <= jend; j++)
      for (int k = kstr + 1; k <= kend - 1; k++)
        for (int i = istr; i <= iend; i++) {
          R aa = K(1.0) / (A3(ss, i, j, k) - A3(rr, i, j, k) * A3(ss, i, j, k - 1));
          A3(ss, i, j, k) = A3(tt, i, j, k) * aa;
          A3(ff, i, j, k) = (A3(ff, i, j, k) - A3(rr, i, j, k) * A3(ff, i, j, k - 1)) * aa;
        }
  }
  if (kstr + 1 <= kend) {
    for (int j = jstr; j <= jend; j++)
      for (int i = istr; i <= iend; i++) {
        R aa = K(1.0) / (A3(ss, i, j, kend) - A3(rr, i, j, kend) * A3(ss, i, j, kem1));
        A3(ff, i, j, kend) = (A3(ff, i, j, kend) - A3(rr, i, j, kend) * A3(ff, i, j, kem1)) * aa;
      }
    OMP_FOR
    for (int j = jstr; j <= jend; j++)
      for (int k = kend - 1; k >= kstr; k--)
        for (int i = istr; i <= iend; i++)
          A3(ff, i, j, k) = A3(ff, i, j, k) - A3(ss, i, j, k) * A3(ff, i, j, k + 1);
  }
}

/* Lateral fills after a velocity/scalar update at one rank: bcycle on
 * periodic sides; on open sides (option 7) the radiative fills of
 * cress_obc.c and the corner average bc4news. */

/* s_stepuv, stepuv.f90:266-280 radiative fills rbcu/rbcv (open sides),
 * :288-310 update, then bcycle (:390-394, periodic sides), bc4news
 * (:400-401, open sides), lbcu/lbcv (no-ops for option 7 with advopt<=3,
 * lbcu.f90:159-200), vbcu/vbcv (:423-424).  cp: the phase speeds (NULL on a
 * periodic domain). */
void or_stepuv(const or_cfg *c, R dts, const R *rst8u, const R *rst8v, const R *ufrc,
               const R *vfrc, const R *usml, const R *vsml, R *uf, R *vf, R *cp)
{
  DECL_DIMS(c);
  R *cx[5], *cy[5];
  if (cp) {
    or_cp_split(c, cp, cx, cy);
    or_rbcu(c, cx[0], cy[0], uf);
    or_rbcv(c, cx[1], cy[1], vf);
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(uf, i, j, k) = A3(uf, i, j, k) + dts * (A3(ufrc, i, j, k) + A3(usml, i, j, k)) / A3(rst8u, i, j, k);
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(vf, i, j, k) = A3(vf, i, j, k) + dts * (A3(vfrc, i, j, k) + A3(vsml, i, j, k)) / A3(rst8v, i, j, k);
  }
  or_bcycle(c, 3, 1, ni - 2, ni, 2, 1, nj - 2, nj - 1, nk, uf);
  or_bcycle(c, 2, 1, ni - 2, ni - 1, 3, 1, nj - 2, nj, nk, vf);
  or_bc4news(c, 1, ni, 1, nj - 1, nk, uf);
  or_bc4news(c, 1, ni - 1, 1, nj, nk, vf);
  or_vbcu(c, uf);
  or_vbcv(c, vf);
}

/* s_stepwi, stepwi.f90:308-549 (buyopt 0/1, impopt=1, flat terrain).
 * fw: explicit w forcing on entry, then reused as the coefficient scratch
 * (the reference's aliasing is kept); wc: contravariant w on exit;
 * tmp1..tmp3: sub/diag/super diagonals. */
void or_stepwi(const or_cfg *c, R dts, const R *jcb, const R *jcb8w, const R *rbr,
               const R *rst, const R *rst8w, const R *rcsq, R *fw, R *wf, R *wc,
               R *tmp1, R *tmp2, R *tmp3, R *cp)
{
  DECL_DIMS(c);
  if (cp) {   /* rbcw on open sides (stepwi.f90:286-296) */
    R *cx[5], *cy[5];
    or_cp_split(c, cp, cx, cy);
    or_rbcw(c, cx[2], cy[2], wf);
  }
  const R dziv = (R)1 / (R)c->dz;
  const R weicoe = (R)c->weicoe;
  const R g05 = K(0.5) * GRAV;
  const R sbsqzi = dts * dts * weicoe * weicoe * dziv;
  const R sbsqg5 = K(0.5) * GRAV * dts * dts * weicoe * weicoe;
  OMP_FOR
  for (int k = 3; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(wf, i, j, k) = A3(wf, i, j, k) + dts * A3(fw, i, j, k) / A3(rst8w, i, j, k);
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        A3(tmp1, i, j, k) = g05 * A3(rbr, i, j, k) / A3(rcsq, i, j, k);
        A3(tmp2, i, j, k) = dziv / A3(jcb, i, j, k);
      }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        A3(fw, i, j, k) = A3(tmp1, i, j, k) + A3(tmp2, i, j, k);
        A3(wc, i, j, k) = A3(tmp1, i, j, k) - A3(tmp2, i, j, k);
      }
  OMP_FOR
  for (int k = 3; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        if (c->buyopt == 0) {
          R a = sbsqzi / A3(rst8w, i, j, k);
          R mm = a * A3(rcsq, i, j, k - 1);
          R nn = a * A3(rcsq, i, j, k);
          A3(tmp1, i, j, k) = -mm * A3(fw, i, j, k - 1);
          A3(tmp2, i, j, k) = K(1.0) + (nn * A3(fw, i, j, k) - mm * A3(wc, i, j, k - 1));
          A3(tmp3, i, j, k) = nn * A3(wc, i, j, k);
        } else {
          R rstiv = K(1.0) / A3(rst8w, i, j, k);
          R a = rstiv * sbsqzi;
          R b = rstiv * sbsqg5;
          R mm = b * A3(rst, i, j, k - 1) - a * A3(rcsq, i, j, k - 1);
          R nn = b * A3(rst, i, j, k) + a * A3(rcsq, i, j, k);
          A3(tmp1, i, j, k) = mm * A3(fw, i, j, k - 1);
          A3(tmp2, i, j, k) = K(1.0) + (mm * A3(wc, i, j, k - 1) + nn * A3(fw, i, j, k));
          A3(tmp3, i, j, k) = nn * A3(wc, i, j, k);
        }
      }
  /* s_bbcw (:426): adds .25*aa*(j31/j32 terms) = .25*aa*0 at k=3 on flat
   * terrain; restated literally so that -0/+0 and NaN behaviour match. */
  for (int j = 2; j <= nj - 2; j++)
    for (int i = 2; i <= ni - 2; i++)
      A3(wf, i, j, 3) = A3(wf, i, j, 3) + K(0.25) * A3(tmp1, i, j, 3) * (K(0.0) + K(0.0));
  or_gaussel(c, 2, ni - 2, 2, nj - 2, 3, nk - 2, tmp1, tmp2, tmp3, wf);
  or_bcycle(c, 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk, wf);
  or_bc4news(c, 1, ni - 1, 1, nj - 1, nk, wf);
  or_phy2cnt(c, c->sthopt, jcb8w, wf, wc);   /* stepwi.f90:527-529, idsthopt */
  or_vbcw(c, jcb8w, wc, wf);
}

/* s_steppi, steppi.f90: rbcss on open sides, :211-223 update, bcycle
 * (:293), bc4news, lbcs (a no-op for option 7 with advopt<=3), vbcp (:316). */
void or_steppi(const or_cfg *c, R dts, const R *jcb, const R *fp, R *ppf, R *cp)
{
  DECL_DIMS(c);
  if (cp) {
    R *cx[5], *cy[5];
    or_cp_split(c, cp, cx, cy);
    or_rbcss(c, cx[3], cy[3], ppf);
  }
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(ppf, i, j, k) = A3(ppf, i, j, k) + dts * A3(fp, i, j, k) / A3(jcb, i, j, k);
  or_bcycle(c, 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk, ppf);
  or_bc4news(c, 1, ni - 1, 1, nj - 1, nk, ppf);
  or_vbcp(c, ppf);
}

/* ------------------------------------------------------------------------ */
/* Step drivers                                                              */
/* ------------------------------------------------------------------------ */

/* Model state for the oracle driver.  Levels: 0 past (*p), 1 present, 2
 * future (*f), as in intgdrv.  q holds the 6 moisture species stacked
 * (qv,qc,qr,qi,qs,qg) and nci the cloud-ice number; the reference's
 * qwtr(:,:,:,1:2)=(qc,qr) and qice(:,:,:,1:3)=(qi,qs,qg) are the slices
 * q+1n and q+3n. */
enum { NQ = 6, QV = 0, QC = 1, QR = 2, QI = 3, QS = 4, QG = 5 };

typedef struct {
  /* base state and metrics, real(0:ni+1,0:nj+1,1:nk) */
  R *jcb, *jcb8u, *jcb8v, *jcb8w, *ubr, *vbr, *pbr, *ptbr, *qvbr, *rbr;
  R *rst, *rst8u, *rst8v, *rst8w, *rcsq;
  R *fc;    /* (0:ni+1,0:nj+1,1:2) */
  R *rbct;  /* (1:ni,1:nj,1:nk,1:2) */
  /* prognostics, three levels */
  R *u[3], *v[3], *w[3], *pp[3], *ptp[3], *q[3], *nci[3];
  /* surface precipitation (0:ni+1,0:nj+1,1:2) x (qc,qr) and x (qi,qs,qg):
   * index 1 rate, 2 accumulation (upwqp.f90:199-202) */
  R *prwtr, *price;
  /* scratch: 3-D fields; at least NSCR of them */
  R *scr;
  /* open-boundary phase speeds (or_cp_size floats), NULL when periodic */
  R *cp;
} or_state;

#define NSCR 24

/* forcedrv (forcedrv.f90:548-619) for the supported options:
 * forceuvw (forceuvw.f90:434-626), forcep (forcep.f90:254-308),
 * forces -> forcept/forceqv/forceq.  Forcings land in ufrc,vfrc,wfrc,pfrc,
 * ptfrc,qvfrc and, as in the reference, in the future q/nci levels
 * (fdmdrv passes qwtrf/qicef/nicef as the forcing arrays). */
static void or_forcedrv(const or_cfg *c, or_state *s, R *qall, R *ufrc, R *vfrc,
                        R *wfrc, R *pfrc, R *ptfrc, R *qvfrc, R *wc)
{
  DECL_DIMS(c);
  const size_t n = (size_t)nfield(c);
  R *t[8];
  for (int i = 0; i < 8; i++) t[i] = s->scr + (size_t)(16 + i) * n;
  R *rstxu = s->scr + 13 * n, *rstxv = s->scr + 14 * n, *rstxwc = s->scr + 15 * n;
  const int P = 0, C = 1, F = 2;
  /* forceuvw */
  memset(ufrc, 0, n * sizeof(R));
  memset(vfrc, 0, n * sizeof(R));
  memset(wfrc, 0, n * sizeof(R));
  or_phy2cnt(c, c->sthopt, s->jcb8w, s->w[C], wc);
  const int sm = c->smtopt % 10 == 2 || c->smtopt % 10 == 3;
  /* smoothing of the past level (forceuvw.f90:467-476, forcep.f90:273-280,
   * forcept.f90:362-370, forceqv.f90:308-316, forceq.f90:413-420 ...) */
  if (sm)
    or_smoo4uvw(c, s->jcb8u, s->jcb8v, s->jcb8w, s->ubr, s->vbr, s->rst8u, s->rst8v, s->rst8w,
                s->u[P], s->v[P], s->w[P], ufrc, vfrc, wfrc, t[0], t[1], t[2], t[3]);
  if (c->vspopt >= 1)
    or_vspuvw(c, s->ubr, s->vbr, s->rst8u, s->rst8v, s->rst8w, s->u[P], s->v[P], s->w[P],
              s->rbct, ufrc, vfrc, wfrc, t[0]);
  if (c->coropt == 1) or_coriuv(c, s->fc, s->rst, s->u[C], s->v[C], ufrc, vfrc, t[0]);
  if (c->buyopt == 1)
    or_buoywb(c, s->ptbr, s->qvbr, s->rst, s->ptp[C], s->q[C], qall, wfrc, t[0], t[1]);
  or_rstuvwc(c, s->rst8u, s->rst8v, s->rst8w, s->u[C], s->v[C], wc, rstxu, rstxv, rstxwc);
  or_advuvw(c, rstxu, rstxv, rstxwc, s->u[C], s->v[C], s->w[C], ufrc, vfrc, wfrc,
            t[0], t[1], t[2], t[3], t[4]);
  /* forcep */
  or_advp(c, s->jcb8u, s->jcb8v, s->jcb8w, s->u[C], s->v[C], wc, s->pp[C], pfrc,
          t[0], t[1], t[2], t[3], t[4], t[5], t[6]);
  /* no smoothing of p': s_forcedrv passes the constant-zero option id as
   * forcep's smtopt (forcedrv.f90:598, idzeropt) */
  if (c->vspopt >= 1 && c->vspvar[3] == 'o') or_vsps(c, s->jcb, s->pp[P], NULL, s->rbct, pfrc);
  /* forcept */
  or_advs(c, rstxu, rstxv, rstxwc, s->ptp[C], ptfrc, t[0], t[1], t[2], t[3]);
  if (sm) or_smoo4s(c, 0, s->rbr, NULL, s->ptp[P], ptfrc, t[0], t[1], t[2], t[3]);
  if (c->vspopt >= 1 && c->vspvar[4] == 'o') or_vsps(c, s->rst, s->ptp[P], NULL, s->rbct, ptfrc);
  or_advbspt(c, s->ptbr, s->rbr, s->w[C], ptfrc, t[0]);
  /* forceqv */
  or_advs(c, rstxu, rstxv, rstxwc, s->q[C], qvfrc, t[0], t[1], t[2], t[3]);
  if (sm) or_smoo4s(c, 1, s->rbr, s->qvbr, s->q[P], qvfrc, t[0], t[1], t[2], t[3]);
  if (c->vspopt >= 1 && c->vspvar[5] == 'o')
    or_vsps(c, s->rst, s->q[P], s->qvbr, s->rbct, qvfrc);
  /* forceq (abs(cphopt)>=1, forceq.f90:384-388): qc, qr, qi, qs, qg, nci */
  if (c->cphopt == 0) return;
  for (int sp = QC; sp <= QG; sp++) {
    or_advs(c, rstxu, rstxv, rstxwc, s->q[C] + sp * n, s->q[F] + sp * n, t[0], t[1], t[2], t[3]);
    if (sm) or_smoo4s(c, 0, s->rbr, NULL, s->q[P] + sp * n, s->q[F] + sp * n, t[0], t[1], t[2], t[3]);
    if (c->vspopt >= 1 && c->vspvar[6] == 'o')
      or_vsps(c, s->rst, s->q[P] + sp * n, NULL, s->rbct, s->q[F] + sp * n);
  }
  or_advs(c, rstxu, rstxv, rstxwc, s->nci[C], s->nci[F], t[0], t[1], t[2], t[3]);
  if (sm) or_smoo4s(c, 0, s->rbr, NULL, s->nci[P], s->nci[F], t[0], t[1], t[2], t[3]);
  if (c->vspopt >= 1 && c->vspvar[6] == 'o') or_vsps(c, s->rst, s->nci[P], NULL, s->rbct, s->nci[F]);
}

/* s_smalldrv, smalldrv.f90:343-398 with s_hevi, hevi.f90:339-427: nsstp
 * acoustic substeps on the future level (uf,vf,wf,ppf). */
/* One acoustic substep, s_hevi (hevi.f90:339-427): pgrad, stepuv,
 * diverpih, diverpiv, advbspi 'back', pgradiv, buoywsi, stepwi, diverpiv,
 * advbspi 'fore', steppi, with tmp[1..6] as hevi's work arrays. */
static void or_substep(const or_cfg *c, R dts, const R *jcb, const R *jcb8u, const R *jcb8v,
                       const R *jcb8w, const R *rbr, const R *rst, const R *rst8u, const R *rst8v,
                       const R *rst8w, const R *rcsq, const R *ufrc, const R *vfrc, const R *wfrc,
                       const R *pfrc, R *uf, R *vf, R *wf, R *wc, R *ppf, R *const *tmp, R *cp)
{
  or_pgrad(c, dts, jcb, rst8u, rst8v, rst8w, uf, vf, wc, ppf, tmp[1], tmp[2], tmp[3], tmp[4], tmp[5], tmp[6]);
  or_stepuv(c, dts, rst8u, rst8v, ufrc, vfrc, tmp[1], tmp[2], uf, vf, cp);
  or_diverpih(c, jcb8u, jcb8v, rcsq, uf, vf, tmp[2], tmp[4], tmp[5]);
  or_diverpiv(c, rcsq, wf, tmp[4]);
  or_advbspi(c, 0, rst, wf, pfrc, tmp[2], tmp[4], tmp[1]);
  or_pgradiv(c, dts, jcb, wfrc, tmp[1], tmp[3], tmp[4]);
  if (c->buyopt == 1) or_buoywsi(c, dts, rbr, rst, rcsq, ppf, tmp[1], tmp[3], tmp[4]);
  or_stepwi(c, dts, jcb, jcb8w, rbr, rst, rst8w, rcsq, tmp[3], wf, wc, tmp[4], tmp[5], tmp[6], cp);
  or_diverpiv(c, rcsq, wf, tmp[4]);
  or_advbspi(c, 1, rst, wf, pfrc, tmp[3], tmp[4], tmp[1]);
  or_steppi(c, dts, jcb, tmp[1], ppf, cp);
}

/* s_hevi for one substep on periodic sides, with caller-owned work arrays
 * t1..t6 (the tier-2 entry point's counterpart). */
void or_hevi(const or_cfg *c, R dts, const R *jcb, const R *jcb8u, const R *jcb8v, const R *jcb8w,
             const R *rbr, const R *rst, const R *rst8u, const R *rst8v, const R *rst8w, const R *rcsq,
             const R *ufrc, const R *vfrc, const R *wfrc, const R *pfrc, R *uf, R *vf, R *wf, R *wc,
             R *ppf, R *t1, R *t2, R *t3, R *t4, R *t5, R *t6)
{
  R *const tmp[7] = {NULL, t1, t2, t3, t4, t5, t6};
  or_substep(c, dts, jcb, jcb8u, jcb8v, jcb8w, rbr, rst, rst8u, rst8v, rst8w, rcsq, ufrc, vfrc, wfrc, pfrc, uf,
             vf, wf, wc, ppf, tmp, NULL);
}

static void or_smalldrv(const or_cfg *c, or_state *s, int nsstp, R dts,
                        const R *ufrc, const R *vfrc, const R *wfrc, const R *pfrc, R *wc)
{
  const size_t n = (size_t)nfield(c);
  R *tmp[7];
  for (int i = 1; i <= 6; i++) tmp[i] = s->scr + (size_t)(16 + i) * n;
  const int P = 0, F = 2;
  memcpy(s->u[F], s->u[P], n * sizeof(R));
  memcpy(s->v[F], s->v[P], n * sizeof(R));
  memcpy(s->w[F], s->w[P], n * sizeof(R));
  memcpy(s->pp[F], s->pp[P], n * sizeof(R));
  or_phy2cnt(c, c->sthopt, s->jcb8w, s->w[F], wc);
  for (int isstp = 1; isstp <= nsstp; isstp++)
    or_substep(c, dts, s->jcb, s->jcb8u, s->jcb8v, s->jcb8w, s->rbr, s->rst, s->rst8u, s->rst8v, s->rst8w,
               s->rcsq, ufrc, vfrc, wfrc, pfrc, s->u[F], s->v[F], s->w[F], wc, s->pp[F], tmp, s->cp);
}

/* s_steps, steps.f90:520-760 update, the radiative fills on open sides
 * (rbcpt/rbcqv/rbcq/rbcs, :1024-1230; lbcs is a no-op for option 7 with
 * advopt<=3), bcycle (:3979-4171, periodic sides), bc4news (:4182-4351,
 * open sides), vbcs (tail). */
static void or_steps(const or_cfg *c, or_state *s, R dtb, const R *ptfrc, const R *qvfrc, R *dtdrst)
{
  DECL_DIMS(c);
  const size_t n = (size_t)nfield(c);
  const int P = 0, F = 2;
  const R dtb2 = K(2.0) * dtb;
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) A3(dtdrst, i, j, k) = dtb2 / A3(s->rst, i, j, k);
  OMP_FOR
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        size_t p = (size_t)i + NX * ((size_t)j + NY * ((size_t)k - 1));
        R d = dtdrst[p];
        s->ptp[F][p] = s->ptp[P][p] + ptfrc[p] * d;
        R a = s->q[P][p] + qvfrc[p] * d;
        s->q[F][p] = a > K(0.0) ? a : K(0.0);
        if (c->cphopt == 0) continue;
        for (int sp = QC; sp <= QG; sp++) {
          R b = s->q[P][p + sp * n] + s->q[F][p + sp * n] * d;
          s->q[F][p + sp * n] = b > K(0.0) ? b : K(0.0);
        }
        s->nci[F][p] = s->nci[P][p] + s->nci[F][p] * d;
      }
  const int nq = c->cphopt == 0 ? 1 : NQ;
  if (s->cp) {
    R *cx[5], *cy[5];
    or_cp_split(c, s->cp, cx, cy);
    const int C = 1;
    or_rbclf(c, 0, cx[4], cy[4], s->ptp[C], s->ptp[P], s->ptp[F]);
    for (int sp = 0; sp < nq; sp++)
      or_rbclf(c, 1, cx[4], cy[4], s->q[C] + sp * n, s->q[P] + sp * n, s->q[F] + sp * n);
    if (c->cphopt) or_rbclf(c, 0, cx[4], cy[4], s->nci[C], s->nci[P], s->nci[F]);
  }
  or_bcycle(c, 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk, s->ptp[F]);
  for (int sp = 0; sp < nq; sp++) or_bcycle(c, 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk, s->q[F] + sp * n);
  if (c->cphopt) or_bcycle(c, 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk, s->nci[F]);
  or_bc4news(c, 1, ni - 1, 1, nj - 1, nk, s->ptp[F]);
  for (int sp = 0; sp < nq; sp++) or_bc4news(c, 1, ni - 1, 1, nj - 1, nk, s->q[F] + sp * n);
  if (c->cphopt) or_bc4news(c, 1, ni - 1, 1, nj - 1, nk, s->nci[F]);
  or_vbcs(c, s->ptp[F]);
  for (int sp = 0; sp < nq; sp++) or_vbcs(c, s->q[F] + sp * n);
  if (c->cphopt) or_vbcs(c, s->nci[F]);
}

/* s_steps on given arrays (the tier-2 entry point's counterpart): qp, qf
 * hold qv, qc, qr, qi, qs, qg one field after another; qf's species and
 * ncif hold the forcings on entry; periodic sides. */
void or_steps_a(const or_cfg *c, R dtb, R *rst, R *ptpp, const R *ptfrc, R *qp, const R *qvfrc, R *ncip,
                R *ptpf, R *qf, R *ncif, R *dtdrst)
{
  or_state st;
  memset(&st, 0, sizeof st);
  st.rst = rst;
  st.ptp[0] = ptpp;
  st.ptp[2] = ptpf;
  st.q[0] = qp;
  st.q[2] = qf;
  st.nci[0] = ncip;
  st.nci[2] = ncif;
  or_steps(c, &st, dtb, ptfrc, qvfrc, dtdrst);
}

/* The TKE lines of s_steps for tubopt>=2, one rank with periodic sides:
 * the leapfrog clipped at zero (steps.f90:997-1012; dtdrst = 2 dtb / rst
 * as above), then bcycle, bc4news and vbcs (steps.f90:4170-4172,
 * 4350-4352, 4493-4495).  tkef holds the forcing on entry. */
void or_stepstke(const or_cfg *c, R dtb, const R *rst, const R *tkep, R *tkef)
{
  DECL_DIMS(c);
  const R dtb2 = K(2.0) * dtb;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        const R d = dtb2 / A3(rst, i, j, k);
        const R a = A3(tkep, i, j, k) + A3(tkef, i, j, k) * d;
        A3(tkef, i, j, k) = a > K(0.0) ? a : K(0.0);
      }
  or_bcycle(c, 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk, tkef);
  or_bc4news(c, 1, ni - 1, 1, nj - 1, nk, tkef);
  or_vbcs(c, tkef);
}

/* s_timeflt, timeflt.f90:335-560: Asselin filter on the present level. */
static void or_asselin(const or_cfg *c, R *x, const R *xp, const R *xf, int ilo, int ihi,
                       int jhi, int khi)
{
  DECL_DIMS(c);
  const R filcoe = (R)c->filcoe;
  const R m1fc2 = K(1.0) - K(2.0) * filcoe;
  OMP_FOR
  for (int k = 1; k <= khi; k++)
    for (int j = 1; j <= jhi; j++)
      for (int i = ilo; i <= ihi; i++)
        A3(x, i, j, k) = m1fc2 * A3(x, i, j, k) + filcoe * (A3(xf, i, j, k) + A3(xp, i, j, k));
}

void or_timeflt(const or_cfg *c, or_state *s)
{
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  const size_t n = (size_t)nfield(c);
  or_asselin(c, s->u[1], s->u[0], s->u[2], 1, ni, nj - 1, nk - 1);
  or_asselin(c, s->v[1], s->v[0], s->v[2], 1, ni - 1, nj, nk - 1);
  or_asselin(c, s->w[1], s->w[0], s->w[2], 1, ni - 1, nj - 1, nk);
  or_asselin(c, s->pp[1], s->pp[0], s->pp[2], 1, ni - 1, nj - 1, nk - 1);
  or_asselin(c, s->ptp[1], s->ptp[0], s->ptp[2], 1, ni - 1, nj - 1, nk - 1);
  const int nq = c->cphopt == 0 ? 1 : NQ;
  for (int sp = 0; sp < nq; sp++)
    or_asselin(c, s->q[1] + sp * n, s->q[0] + sp * n, s->q[2] + sp * n, 1, ni - 1, nj - 1, nk - 1);
  if (c->cphopt) or_asselin(c, s->nci[1], s->nci[0], s->nci[2], 1, ni - 1, nj - 1, nk - 1);
}

/* s_shift2nd at one rank (shift2nd.f90:4274-4302): second halo row for the
 * 4th-order stencils (periodic sides); on open sides only the corners of
 * the outer ring, bc4news (:4590-4700). */
void or_shift2nd(const or_cfg *c, or_state *s, int lv)
{
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  const size_t n = (size_t)nfield(c);
  /* shift2nd.f90:266-268: nothing to do unless a 4th-order stencil reads
   * the second halo (advopt>=2 or smtopt 2/3; tubopt=0, no option-6 side) */
  if (!(c->advopt >= 2 || c->smtopt % 10 == 2 || c->smtopt % 10 == 3)) return;
  or_bcycle(c, 4, 0, ni - 3, ni + 1, 3, 0, nj - 3, nj, nk, s->u[lv]);
  or_bcycle(c, 3, 0, ni - 3, ni, 4, 0, nj - 3, nj + 1, nk, s->v[lv]);
  or_bcycle(c, 3, 0, ni - 3, ni, 3, 0, nj - 3, nj, nk, s->w[lv]);
  or_bcycle(c, 3, 0, ni - 3, ni, 3, 0, nj - 3, nj, nk, s->pp[lv]);
  or_bcycle(c, 3, 0, ni - 3, ni, 3, 0, nj - 3, nj, nk, s->ptp[lv]);
  const int nq = c->cphopt == 0 ? 1 : NQ;
  for (int sp = 0; sp < nq; sp++) or_bcycle(c, 3, 0, ni - 3, ni, 3, 0, nj - 3, nj, nk, s->q[lv] + sp * n);
  if (c->cphopt) or_bcycle(c, 3, 0, ni - 3, ni, 3, 0, nj - 3, nj, nk, s->nci[lv]);
  or_bc4news(c, 0, ni + 1, 0, nj, nk, s->u[lv]);
  or_bc4news(c, 0, ni, 0, nj + 1, nk, s->v[lv]);
  or_bc4news(c, 0, ni, 0, nj, nk, s->w[lv]);
  or_bc4news(c, 0, ni, 0, nj, nk, s->pp[lv]);
  or_bc4news(c, 0, ni, 0, nj, nk, s->ptp[lv]);
  for (int sp = 0; sp < nq; sp++) or_bc4news(c, 0, ni, 0, nj, nk, s->q[lv] + sp * n);
  if (c->cphopt) or_bc4news(c, 0, ni, 0, nj, nk, s->nci[lv]);
}

/* s_swp2nxt, swp2nxt.f90:354-...: past <- present <- future over the
 * reference's copy ranges (u: i=iwest..ni+1-ieast, j=jsouth..nj-jnorth,
 * k=1..nk-1; v: i=iwest..ni-ieast, j=jsouth..nj+1-jnorth; w and scalars:
 * i=iwest..ni-ieast, j=jsouth..nj-jnorth, k=1..nk (w) / nk-1). */
static void or_rot(const or_cfg *c, R *xp, R *x, const R *xf, int ihi, int jhi, int khi)
{
  DECL_DIMS(c);
  for (int k = 1; k <= khi; k++)
    for (int j = c->jsouth; j <= jhi; j++)
      for (int i = c->iwest; i <= ihi; i++) {
        A3(xp, i, j, k) = A3(x, i, j, k);
        A3(x, i, j, k) = A3(xf, i, j, k);
      }
}

void or_swp2nxt(const or_cfg *c, or_state *s)
{
  const int ni = c->ni, nj = c->nj, nk = c->nk;
  const int ie = c->ieast, jn = c->jnorth;
  const size_t n = (size_t)nfield(c);
  or_rot(c, s->u[0], s->u[1], s->u[2], ni + 1 - ie, nj - jn, nk - 1);
  or_rot(c, s->v[0], s->v[1], s->v[2], ni - ie, nj + 1 - jn, nk - 1);
  or_rot(c, s->w[0], s->w[1], s->w[2], ni - ie, nj - jn, nk);
  or_rot(c, s->pp[0], s->pp[1], s->pp[2], ni - ie, nj - jn, nk - 1);
  or_rot(c, s->ptp[0], s->ptp[1], s->ptp[2], ni - ie, nj - jn, nk - 1);
  const int nq = c->cphopt == 0 ? 1 : NQ;
  for (int sp = 0; sp < nq; sp++)
    or_rot(c, s->q[0] + sp * n, s->q[1] + sp * n, s->q[2] + sp * n, ni - ie, nj - jn, nk - 1);
  if (c->cphopt) or_rot(c, s->nci[0], s->nci[1], s->nci[2], ni - ie, nj - jn, nk - 1);
}

/* Configurations this restatement covers (the plain-path options). */
int or_cfg_check(const or_cfg *c)
{
  if (c->ni < 7 || c->nj < 7 || c->nk < 7) return -1;
  const int per = c->wbc == -1 && c->ebc == -1 && c->sbc == -1 && c->nbc == -1;
  const int opn = c->wbc == 7 && c->ebc == 7 && c->sbc == 7 && c->nbc == 7;
  if (!per && !opn) return -2;
  if (c->bbc != 2 && c->bbc != 3) return -3;
  if (c->tbc != 2 && c->tbc != 3) return -3;
  if (c->advopt < 1 || c->advopt > 3) return -4;
  if (c->impopt != 1 || c->gwmopt != 0 || c->divopt < 0 || c->divopt > 1) return -5;
  if (c->cphopt != 0 && c->cphopt != 2) return -6;
  if (c->haiopt != 0) return -6;
  if (c->vspopt != 0 && c->vspopt != 2) return -7;
  if (c->smtopt != 0 && c->smtopt != 2 && c->smtopt != 3) return -9;
  return 0;
}

#include "cress_micro.c"

typedef void (*or_micro_fn)(const or_cfg *, or_state *, R dtb, R *qallp);
static or_micro_fn g_micro = micro_hook;
static or_micro_fn g_satadj = satadj_hook;

/* One large step: s_intgdrv (intgdrv.f90:832-1133) minus I/O, with the
 * time bookkeeping of s_timeint (timeint.f90:160-190): the first step
 * (first!=0, ctime=0) is a forward step with dtb=dtbig/2, dts=dtsml/2 and no
 * Asselin filter (intgdrv.f90:968). */
int or_large_step(const or_cfg *c, or_state *s, int first, double dtbig, double dtsml, int nsstp)
{
  if (or_cfg_check(c)) return or_cfg_check(c);
  if (or_open(c) && !s->cp) return -8;
  const size_t n = (size_t)nfield(c);
  R *ufrc = s->scr, *vfrc = s->scr + n, *wfrc = s->scr + 2 * n, *pfrc = s->scr + 3 * n;
  R *ptfrc = s->scr + 4 * n, *qvfrc = s->scr + 5 * n, *wc = s->scr + 6 * n;
  R *qall = s->scr + 7 * n, *qallp = s->scr + 8 * n, *dtdrst = s->scr + 9 * n;
  const R dtb = first ? K(0.5) * (R)dtbig : (R)dtbig;
  const R dts = first ? K(0.5) * (R)dtsml : (R)dtsml;
  if (c->cphopt) {
    or_totalqwi(c, s->q[1] + n, s->q[1] + 3 * n, qall);
    or_totalqwi(c, s->q[0] + n, s->q[0] + 3 * n, qallp);
  }
  or_forcedrv(c, s, qall, ufrc, vfrc, wfrc, pfrc, ptfrc, qvfrc, wc);
  or_smalldrv(c, s, nsstp, dts, ufrc, vfrc, wfrc, pfrc, wc);
  or_steps(c, s, dtb, ptfrc, qvfrc, dtdrst);
  if (c->cphopt == 2 && g_micro) g_micro(c, s, dtb, qallp);
  if (!first) or_timeflt(c, s);
  if (c->cphopt == 2 && g_satadj) g_satadj(c, s, dtb, NULL);
  or_shift2nd(c, s, 1);
  or_shift2nd(c, s, 2);
  if (s->cp) or_phasev(c, dtb, dts, s->cp);
  or_swp2nxt(c, s);
  return 0;
}
#include "cress_turb.c"
