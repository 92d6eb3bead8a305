/* Open lateral boundaries (wbc=ebc=sbc=nbc=7: radiation with a constant
 * phase speed) -- TEST INFRASTRUCTURE ONLY.  Included by cress_oracle.c.
 *
 * Restated for advopt<=3, exbopt=0, mfcopt=0, nggopt=lspopt=0, vspopt!=1 and
 * lbcvar='x' everywhere (no lateral damping: dmpdt=0), i.e. the reference's
 * plain open-boundary path.  The damping terms are zero there and are kept
 * only where they change the operation order of what remains (p - 0*p = p
 * exactly, so none does).
 *
 * Phase-speed arrays keep the reference's shapes: *cpx(1:nj,1:nk,1:2) for
 * the west (1) / east (2) sides, *cpy(1:ni,1:nk,1:2) for south / north.
 * The reference holds one pair per prognostic; for option 7 every scalar's
 * pair is the same function of (dtb, dts) (phvs.f90:375-382 ignores the
 * field), so the large-step scalars (theta', qv, the q species, nci) share
 * one pair here: ptcpx = qvcpx = qwcpx(..,n) = qicpx(..,n) = nicpx. */

#define CPX(a, j, k, s) (a)[(size_t)((j) - 1) + (size_t)nj * ((size_t)((k) - 1) + (size_t)nk * (size_t)((s) - 1))]
#define CPY(a, i, k, s) (a)[(size_t)((i) - 1) + (size_t)ni * ((size_t)((k) - 1) + (size_t)nk * (size_t)((s) - 1))]

static int or_open(const or_cfg *c) { return c->wbc == 7; }

/* Offsets of the phase-speed arrays in one block:
 * u(x,y) v(x,y) w(x,y) p(x,y) s(x,y). */
size_t or_cp_size(const or_cfg *c) { return (size_t)5 * 2 * (size_t)c->nk * (size_t)(c->ni + c->nj); }
static void or_cp_split(const or_cfg *c, R *cp, R **x, R **y)
{
  const size_t nx = (size_t)2 * c->nk * c->nj, ny = (size_t)2 * c->nk * c->ni;
  for (int v = 0; v < 5; v++) {
    x[v] = cp + (size_t)v * (nx + ny);
    y[v] = x[v] + nx;
  }
}

/* s_phvuvw, phvuvw.f90:334-346 (constants), option-7 branches with the
 * mfcopt=0 clip and the dts/dtb scaling: u :437-493 (west; set and clip
 * k=2..nk-1, scale k=2..nk-2), :578-634 (east), :752-808 (south), north;
 * v :1306-1372 (west, j=1..nj), east, south (i=1..ni-1), north; w (k=2..nk-1
 * throughout, j=1..nj-1 / i=1..ni-1). */
void or_phvuvw(const or_cfg *c, R dtb, R dts, R *ucpx, R *ucpy, R *vcpx, R *vcpy, R *wcpx,
               R *wcpy)
{
  DECL_DIMS(c);
  if (!or_open(c)) return;
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy;
  const R gwave = (R)c->gwave;
  const R gdxdt = gwave * dtb * dxiv, gdydt = gwave * dtb * dyiv;
  const R gdxdtn = -(gwave * dtb * dxiv), gdydtn = -(gwave * dtb * dyiv);
  const R dtsdb = dts / dtb;
  /* u */
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++) CPX(ucpx, j, k, 1) = gdxdtn;
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++) CPX(ucpx, j, k, 1) = fmax(CPX(ucpx, j, k, 1), K(-1.0));
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++) CPX(ucpx, j, k, 1) = CPX(ucpx, j, k, 1) * dtsdb;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++) {
      CPX(ucpx, j, k, 2) = gdxdt;
      CPX(ucpx, j, k, 2) = fmin(CPX(ucpx, j, k, 2), K(1.0));
      CPX(ucpx, j, k, 2) = CPX(ucpx, j, k, 2) * dtsdb;
    }
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 1; i <= ni; i++) {
      CPY(ucpy, i, k, 1) = fmax(gdydtn, K(-1.0)) * dtsdb;
      CPY(ucpy, i, k, 2) = fmin(gdydt, K(1.0)) * dtsdb;
    }
  /* v */
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj; j++) {
      CPX(vcpx, j, k, 1) = fmax(gdxdtn, K(-1.0)) * dtsdb;
      CPX(vcpx, j, k, 2) = fmin(gdxdt, K(1.0)) * dtsdb;
    }
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 1; i <= ni - 1; i++) {
      CPY(vcpy, i, k, 1) = fmax(gdydtn, K(-1.0)) * dtsdb;
      CPY(vcpy, i, k, 2) = fmin(gdydt, K(1.0)) * dtsdb;
    }
  /* w */
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++) {
      CPX(wcpx, j, k, 1) = fmax(gdxdtn, K(-1.0)) * dtsdb;
      CPX(wcpx, j, k, 2) = fmin(gdxdt, K(1.0)) * dtsdb;
    }
  for (int k = 2; k <= nk - 1; k++)
    for (int i = 1; i <= ni - 1; i++) {
      CPY(wcpy, i, k, 1) = fmax(gdydtn, K(-1.0)) * dtsdb;
      CPY(wcpy, i, k, 2) = fmin(gdydt, K(1.0)) * dtsdb;
    }
}

/* s_phvs, phvs.f90:270-290 (constants; dtdvb=dts/dtb for 'sml'), option-7
 * branches :375-460 (west), :538-620 (east), south, north: k=2..nk-2,
 * j=1..nj-1 / i=1..ni-1, clip to [-1,1], scaled by dtdvb only for the
 * acoustic ('sml') pressure -- with advopt<=3 the large-step scalars keep
 * the unscaled speed. */
void or_phvs(const or_cfg *c, int sml, R dtb, R dts, R *scpx, R *scpy)
{
  DECL_DIMS(c);
  if (!or_open(c)) return;
  const R dxiv = (R)1 / (R)c->dx, dyiv = (R)1 / (R)c->dy;
  const R gwave = (R)c->gwave;
  const R gdxdt = gwave * dtb * dxiv, gdydt = gwave * dtb * dyiv;
  const R gdxdtn = -(gwave * dtb * dxiv), gdydtn = -(gwave * dtb * dyiv);
  const R dtdvb = dts / dtb;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++) {
      R w = fmax(gdxdtn, K(-1.0)), e = fmin(gdxdt, K(1.0));
      if (sml) {
        w = w * dtdvb;
        e = e * dtdvb;
      }
      CPX(scpx, j, k, 1) = w;
      CPX(scpx, j, k, 2) = e;
    }
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 1; i <= ni - 1; i++) {
      R s = fmax(gdydtn, K(-1.0)), n = fmin(gdydt, K(1.0));
      if (sml) {
        s = s * dtdvb;
        n = n * dtdvb;
      }
      CPY(scpy, i, k, 1) = s;
      CPY(scpy, i, k, 2) = n;
    }
}

/* s_phasev, phasev.f90:464-500 for gwmopt=0: u,v,w; p ('sml'); the
 * large-step scalars ('big').  Called at the end of a large step
 * (intgdrv.f90:1023-1055); the speeds it leaves are the next step's. */
void or_phasev(const or_cfg *c, R dtb, R dts, R *cp)
{
  R *x[5], *y[5];
  or_cp_split(c, cp, x, y);
  or_phvuvw(c, dtb, dts, x[0], y[0], x[1], y[1], x[2], y[2]);
  or_phvs(c, 1, dtb, dts, x[3], y[3]);
  or_phvs(c, 0, dtb, dts, x[4], y[4]);
}

/* s_rbcu, rbcu.f90:262-400 (ndmpdt=tdmpdt=0): west/east over j=1..nj-1,
 * south/north over i=2..ni-1, k=2..nk-2, in that order. */
void or_rbcu(const or_cfg *c, const R *ucpx, const R *ucpy, R *u)
{
  DECL_DIMS(c);
  if (!or_open(c)) return;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      A3(u, 1, j, k) = A3(u, 1, j, k) - CPX(ucpx, j, k, 1) * (A3(u, 2, j, k) - A3(u, 1, j, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      A3(u, ni, j, k) = A3(u, ni, j, k) + CPX(ucpx, j, k, 2) * (A3(u, ni - 1, j, k) - A3(u, ni, j, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 2; i <= ni - 1; i++)
      A3(u, i, 1, k) = A3(u, i, 1, k) - CPY(ucpy, i, k, 1) * (A3(u, i, 2, k) - A3(u, i, 1, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 2; i <= ni - 1; i++)
      A3(u, i, nj - 1, k) = A3(u, i, nj - 1, k) + CPY(ucpy, i, k, 2) * (A3(u, i, nj - 2, k) - A3(u, i, nj - 1, k));
}

/* s_rbcv, rbcv.f90:262-400: south/north over i=1..ni-1, then west/east over
 * j=2..nj-1, k=2..nk-2. */
void or_rbcv(const or_cfg *c, const R *vcpx, const R *vcpy, R *v)
{
  DECL_DIMS(c);
  if (!or_open(c)) return;
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 1; i <= ni - 1; i++)
      A3(v, i, 1, k) = A3(v, i, 1, k) - CPY(vcpy, i, k, 1) * (A3(v, i, 2, k) - A3(v, i, 1, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 1; i <= ni - 1; i++)
      A3(v, i, nj, k) = A3(v, i, nj, k) + CPY(vcpy, i, k, 2) * (A3(v, i, nj - 1, k) - A3(v, i, nj, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 1; j++)
      A3(v, 1, j, k) = A3(v, 1, j, k) - CPX(vcpx, j, k, 1) * (A3(v, 2, j, k) - A3(v, 1, j, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 1; j++)
      A3(v, ni - 1, j, k) = A3(v, ni - 1, j, k) + CPX(vcpx, j, k, 2) * (A3(v, ni - 2, j, k) - A3(v, ni - 1, j, k));
}

/* Corner-then-edge radiation of a field living on scalar columns
 * (s_rbcw, rbcw.f90:236-470, k=2..nk-1; s_rbcss, rbcss.f90:205-400,
 * k=2..nk-2), dmpdt=0.  Corners first, each from the not yet updated edge
 * values; then west/east over j=2..nj-2 and south/north over i=2..ni-2. */
static void or_rbc_inplace(const or_cfg *c, int khi, const R *cx, const R *cy, R *s)
{
  DECL_DIMS(c);
  if (!or_open(c)) return;
  const int im1 = ni - 1, im2 = ni - 2, jm1 = nj - 1, jm2 = nj - 2;
  for (int k = 2; k <= khi; k++) {
    R we = (A3(s, 2, 1, k) - A3(s, 1, 1, k)) * CPX(cx, 1, k, 1);
    R sn = (A3(s, 1, 2, k) - A3(s, 1, 1, k)) * CPY(cy, 1, k, 1);
    A3(s, 1, 1, k) = A3(s, 1, 1, k) - (we + sn);
  }
  for (int k = 2; k <= khi; k++) {
    R we = (A3(s, im2, 1, k) - A3(s, im1, 1, k)) * CPX(cx, 1, k, 2);
    R sn = (A3(s, im1, 2, k) - A3(s, im1, 1, k)) * CPY(cy, im1, k, 1);
    A3(s, im1, 1, k) = A3(s, im1, 1, k) + (we - sn);
  }
  for (int k = 2; k <= khi; k++) {
    R we = (A3(s, 2, jm1, k) - A3(s, 1, jm1, k)) * CPX(cx, jm1, k, 1);
    R sn = (A3(s, 1, jm2, k) - A3(s, 1, jm1, k)) * CPY(cy, 1, k, 2);
    A3(s, 1, jm1, k) = A3(s, 1, jm1, k) - (we - sn);
  }
  for (int k = 2; k <= khi; k++) {
    R we = (A3(s, im2, jm1, k) - A3(s, im1, jm1, k)) * CPX(cx, jm1, k, 2);
    R sn = (A3(s, im1, jm2, k) - A3(s, im1, jm1, k)) * CPY(cy, im1, k, 2);
    A3(s, im1, jm1, k) = A3(s, im1, jm1, k) + (we + sn);
  }
  for (int k = 2; k <= khi; k++)
    for (int j = 2; j <= nj - 2; j++)
      A3(s, 1, j, k) = A3(s, 1, j, k) - CPX(cx, j, k, 1) * (A3(s, 2, j, k) - A3(s, 1, j, k));
  for (int k = 2; k <= khi; k++)
    for (int j = 2; j <= nj - 2; j++)
      A3(s, im1, j, k) = A3(s, im1, j, k) + CPX(cx, j, k, 2) * (A3(s, im2, j, k) - A3(s, im1, j, k));
  for (int k = 2; k <= khi; k++)
    for (int i = 2; i <= ni - 2; i++)
      A3(s, i, 1, k) = A3(s, i, 1, k) - CPY(cy, i, k, 1) * (A3(s, i, 2, k) - A3(s, i, 1, k));
  for (int k = 2; k <= khi; k++)
    for (int i = 2; i <= ni - 2; i++)
      A3(s, i, jm1, k) = A3(s, i, jm1, k) + CPY(cy, i, k, 2) * (A3(s, i, jm2, k) - A3(s, i, jm1, k));
}

void or_rbcw(const or_cfg *c, const R *wcpx, const R *wcpy, R *w)
{
  or_rbc_inplace(c, c->nk - 1, wcpx, wcpy, w);
}

void or_rbcss(const or_cfg *c, const R *pcpx, const R *pcpy, R *pp)
{
  or_rbc_inplace(c, c->nk - 2, pcpx, pcpy, pp);
}

/* Leapfrog radiation of a large-step scalar, advopt<=3 branch, dmpdt=0:
 * s_rbcpt (rbcpt.f90:252-620), s_rbcqv (rbcqv.f90), s_rbcq (rbcq.f90),
 * s_rbcs (rbcs.f90).  sf on the boundary from the past value sp there and
 * the present value s one cell inside.  Corners first (rw,rs =
 * (s_in - sp)*c/(1 -+ c); sf = sp -+ 2*(rw +- rs)), then the edges with
 * gamma = 2*c/(1 -+ c): west/east j=2..nj-2, south/north i=2..ni-2,
 * k=2..nk-2.  clip0: max(.,0) (rbcqv, rbcq); rbcpt and rbcs do not clip. */
static R or_max0(int clip0, R a) { return clip0 ? (a > K(0.0) ? a : K(0.0)) : a; }

void or_rbclf(const or_cfg *c, int clip0, const R *cx, const R *cy, const R *s, const R *sp, R *sf)
{
  DECL_DIMS(c);
  if (!or_open(c)) return;
  const int im1 = ni - 1, im2 = ni - 2, jm1 = nj - 1, jm2 = nj - 2;
  for (int k = 2; k <= nk - 2; k++) {
    R rw = (A3(s, 2, 1, k) - A3(sp, 1, 1, k)) * CPX(cx, 1, k, 1) / (K(1.0) - CPX(cx, 1, k, 1));
    R rs = (A3(s, 1, 2, k) - A3(sp, 1, 1, k)) * CPY(cy, 1, k, 1) / (K(1.0) - CPY(cy, 1, k, 1));
    A3(sf, 1, 1, k) = or_max0(clip0, A3(sp, 1, 1, k) - K(2.0) * (rw + rs));
  }
  for (int k = 2; k <= nk - 2; k++) {
    R rw = (A3(s, im2, 1, k) - A3(sp, im1, 1, k)) * CPX(cx, 1, k, 2) / (K(1.0) + CPX(cx, 1, k, 2));
    R rs = (A3(s, im1, 2, k) - A3(sp, im1, 1, k)) * CPY(cy, im1, k, 1) / (K(1.0) - CPY(cy, im1, k, 1));
    A3(sf, im1, 1, k) = or_max0(clip0, A3(sp, im1, 1, k) + K(2.0) * (rw - rs));
  }
  for (int k = 2; k <= nk - 2; k++) {
    R rw = (A3(s, 2, jm1, k) - A3(sp, 1, jm1, k)) * CPX(cx, jm1, k, 1) / (K(1.0) - CPX(cx, jm1, k, 1));
    R rs = (A3(s, 1, jm2, k) - A3(sp, 1, jm1, k)) * CPY(cy, 1, k, 2) / (K(1.0) + CPY(cy, 1, k, 2));
    A3(sf, 1, jm1, k) = or_max0(clip0, A3(sp, 1, jm1, k) - K(2.0) * (rw - rs));
  }
  for (int k = 2; k <= nk - 2; k++) {
    R rw = (A3(s, im2, jm1, k) - A3(sp, im1, jm1, k)) * CPX(cx, jm1, k, 2) / (K(1.0) + CPX(cx, jm1, k, 2));
    R rs = (A3(s, im1, jm2, k) - A3(sp, im1, jm1, k)) * CPY(cy, im1, k, 2) / (K(1.0) + CPY(cy, im1, k, 2));
    A3(sf, im1, jm1, k) = or_max0(clip0, A3(sp, im1, jm1, k) + K(2.0) * (rw + rs));
  }
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++) {
      R g = K(2.0) * CPX(cx, j, k, 1) / (K(1.0) - CPX(cx, j, k, 1));
      A3(sf, 1, j, k) = or_max0(clip0, A3(sp, 1, j, k) - g * (A3(s, 2, j, k) - A3(sp, 1, j, k)));
    }
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++) {
      R g = K(2.0) * CPX(cx, j, k, 2) / (K(1.0) + CPX(cx, j, k, 2));
      A3(sf, im1, j, k) = or_max0(clip0, A3(sp, im1, j, k) + g * (A3(s, im2, j, k) - A3(sp, im1, j, k)));
    }
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 2; i <= ni - 2; i++) {
      R g = K(2.0) * CPY(cy, i, k, 1) / (K(1.0) - CPY(cy, i, k, 1));
      A3(sf, i, 1, k) = or_max0(clip0, A3(sp, i, 1, k) - g * (A3(s, i, 2, k) - A3(sp, i, 1, k)));
    }
  for (int k = 2; k <= nk - 2; k++)
    for (int i = 2; i <= ni - 2; i++) {
      R g = K(2.0) * CPY(cy, i, k, 2) / (K(1.0) + CPY(cy, i, k, 2));
      A3(sf, i, jm1, k) = or_max0(clip0, A3(sp, i, jm1, k) + g * (A3(s, i, jm2, k) - A3(sp, i, jm1, k)));
    }
}

/* s_bc4news, bc4news.f90:153-204: the four domain corners of the box
 * isw..ise x jss..jsn become the mean of their x and y neighbours, k=1..kmax,
 * when any side is not periodic. */
void or_bc4news(const or_cfg *c, int isw, int ise, int jss, int jsn, int kmax, R *var)
{
  DECL_DIMS(c);
  if (c->wbc == -1 && c->ebc == -1 && c->sbc == -1 && c->nbc == -1) return;
  for (int k = 1; k <= kmax; k++)
    A3(var, isw, jss, k) = K(0.5) * (A3(var, isw + 1, jss, k) + A3(var, isw, jss + 1, k));
  for (int k = 1; k <= kmax; k++)
    A3(var, ise, jss, k) = K(0.5) * (A3(var, ise - 1, jss, k) + A3(var, ise, jss + 1, k));
  for (int k = 1; k <= kmax; k++)
    A3(var, isw, jsn, k) = K(0.5) * (A3(var, isw + 1, jsn, k) + A3(var, isw, jsn - 1, k));
  for (int k = 1; k <= kmax; k++)
    A3(var, ise, jsn, k) = K(0.5) * (A3(var, ise - 1, jsn, k) + A3(var, ise, jsn - 1, k));
}
