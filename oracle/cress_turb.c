/* The coefficients of the 1.5-order TKE closure (tubopt=2) -- TEST
 * INFRASTRUCTURE ONLY.  Included by cress_oracle.c after cress_micro.c,
 * whose physical constants (C_*) it uses.
 *
 * The first four calls of s_turbdrv (turbdrv.f90:245-276) for flat
 * terrain without map factors (trnopt=0, mfcopt=0), no surface physics
 * (sfcopt=0) and isotropic grids (isoopt=1):
 *   or_bruntv   s_bruntv    bruntv.f90:180-392   N^2 at w points
 *   or_defomten s_defomten  defomten.f90:253-1103 deformation tensor
 *   or_defomssq s_defomssq  defomssq.f90:95-120   its magnitude squared
 *   or_eddyvis  s_eddyvis   eddyvis.f90:216-604   Deardorff eddy viscosity
 * Each writes the index ranges the reference writes and leaves the rest of
 * its outputs as they were.  Fortran's left-to-right evaluation is kept
 * term by term. */

/* s_bruntv: pt, ptv and a are work arrays, t a 2-D one (ni+2)*(nj+2). */
void or_bruntv(const or_cfg *c, int moist, const R *jcb8w, const R *pbr, const R *ptbr, const R *pp,
               const R *ptp, const R *qv, const R *qall, R *nsq8w, R *pt, R *ptv, R *a, R *t)
{
  DECL_DIMS(c);
  const R dziv = K(1.0) / (R)c->dz;
  const R gdzv = C_G * dziv, gdzv05 = K(0.5) * C_G * dziv;
  const R rddvcp = C_RD / C_CP, cwmci = C_CW - C_CI, p0iv = K(1.0) / C_P0;
  const R thresq = (R)c->thresq;
  const R epsav = K(1.6077);
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(pt, i, j, k) = A3(ptbr, i, j, k) + A3(ptp, i, j, k);
  if (!moist || c->cphopt == 0) {
    const R *s = pt;
    if (moist) {
      for (int k = 1; k <= nk - 1; k++)
        for (int j = 1; j <= nj - 1; j++)
          for (int i = 1; i <= ni - 1; i++)
            A3(ptv, i, j, k) = A3(pt, i, j, k) * (K(1.0) + epsav * A3(qv, i, j, k)) / (K(1.0) + A3(qv, i, j, k));
      s = ptv;
    }
    for (int k = 2; k <= nk - 1; k++)
      for (int j = 1; j <= nj - 1; j++)
        for (int i = 1; i <= ni - 1; i++)
          A3(nsq8w, i, j, k) = gdzv * (A3(s, i, j, k) - A3(s, i, j, k - 1)) /
                               (A3(jcb8w, i, j, k) * (A3(ptbr, i, j, k - 1) + A3(ptbr, i, j, k)));
  } else {
    for (int k = 1; k <= nk - 1; k++)
      for (int j = 1; j <= nj - 1; j++)
        for (int i = 1; i <= ni - 1; i++)
          A3(ptv, i, j, k) = A3(pt, i, j, k) * (K(1.0) + epsav * A3(qv, i, j, k)) / (K(1.0) + A3(qv, i, j, k));
    for (int k = 1; k <= nk - 1; k++)
      for (int j = 1; j <= nj - 1; j++)
        for (int i = 1; i <= ni - 1; i++) {
          const size_t p = (size_t)i + NX * (size_t)j;
          t[p] = A3(pt, i, j, k) * EXP(rddvcp * LOG(p0iv * (A3(pbr, i, j, k) + A3(pp, i, j, k))));
          R ak = C_LV0 * EXP((K(.167) + K(3.67e-4) * t[p]) * LOG(C_T0 / t[p]));
          if (t[p] <= C_TLOW) ak = ak + (C_LF0 + cwmci * (t[p] - C_T0));
          const R lhcpt = ak / (C_CP * t[p]);
          A3(pt, i, j, k) = A3(pt, i, j, k) * EXP(lhcpt * A3(qv, i, j, k));
          ak = ak * A3(qv, i, j, k) / (C_RD * t[p]);
          A3(a, i, j, k) = (K(1.0) + ak) / (K(1.0) + C_EPSVA * lhcpt * ak);
        }
    for (int k = 2; k <= nk - 1; k++)
      for (int j = 1; j <= nj - 1; j++)
        for (int i = 1; i <= ni - 1; i++) {
          if (A3(qall, i, j, k) > thresq)
            A3(nsq8w, i, j, k) = gdzv05 * ((A3(qall, i, j, k - 1) - A3(qall, i, j, k)) +
                                           (A3(pt, i, j, k) - A3(pt, i, j, k - 1)) *
                                               (A3(a, i, j, k - 1) + A3(a, i, j, k)) /
                                               (A3(ptbr, i, j, k - 1) + A3(ptbr, i, j, k))) /
                                 A3(jcb8w, i, j, k);
          else
            A3(nsq8w, i, j, k) = gdzv * (A3(ptv, i, j, k) - A3(ptv, i, j, k - 1)) /
                                 (A3(jcb8w, i, j, k) * (A3(ptbr, i, j, k - 1) + A3(ptbr, i, j, k)));
        }
  }
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni - 1; i++) {
      A3(nsq8w, i, j, 1) = A3(nsq8w, i, j, 2);
      A3(nsq8w, i, j, nk) = A3(nsq8w, i, j, nk - 1);
    }
}

/* s_bcten (bcten.f90): the bottom and top values of a tensor component at
 * uw / vw points, mirrored (bc 2) or copied (bc >= 3) from level 3 / nk-2. */
static void or_bcten(const or_cfg *c, R *ten)
{
  DECL_DIMS(c);
  for (int j = 0; j <= nj + 1; j++)
    for (int i = 0; i <= ni + 1; i++) {
      if (c->bbc == 2) A3(ten, i, j, 1) = -A3(ten, i, j, 3);
      else if (c->bbc >= 3) A3(ten, i, j, 1) = A3(ten, i, j, 3);
      if (c->tbc == 2) A3(ten, i, j, nk) = -A3(ten, i, j, nk - 2);
      else if (c->tbc >= 3) A3(ten, i, j, nk) = A3(ten, i, j, nk - 2);
    }
}

/* s_defomten, trnopt=0 and mfcopt=0.  The reference's tmp1 (u*jcb8u,
 * v*jcb8v, w*jcb8w) and tmp3 (inverse mean Jacobians) are formed where
 * used: the same single products and quotients. */
void or_defomten(const or_cfg *c, const R *jcb, const R *jcb8u, const R *jcb8v, const R *jcb8w, const R *u,
                 const R *v, const R *w, R *s11, R *s22, R *s33, R *s12, R *s13, R *s23, R *s31, R *s32)
{
  DECL_DIMS(c);
  const int iw = c->iwest, ie = c->ieast, js = c->jsouth, jn = c->jnorth;
  const R dxiv = K(1.0) / (R)c->dx, dyiv = K(1.0) / (R)c->dy, dziv = K(1.0) / (R)c->dz;
#define UJ(i, j, k) (A3(u, i, j, k) * A3(jcb8u, i, j, k))
#define VJ(i, j, k) (A3(v, i, j, k) * A3(jcb8v, i, j, k))
#define WJ(i, j, k) (A3(w, i, j, k) * A3(jcb8w, i, j, k))
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1 + js; j <= nj - jn; j++)
      for (int i = 1 + iw; i <= ni - ie; i++) {
        const R t3 = K(4.0) / ((A3(jcb, i - 1, j - 1, k) + A3(jcb, i, j, k)) +
                               (A3(jcb, i - 1, j, k) + A3(jcb, i, j - 1, k)));
        A3(s12, i, j, k) = t3 * ((UJ(i, j, k) - UJ(i, j - 1, k)) * dyiv + (VJ(i, j, k) - VJ(i - 1, j, k)) * dxiv);
      }
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        const R jcbiv2 = K(2.0) / A3(jcb, i, j, k);
        A3(s11, i, j, k) = jcbiv2 * ((UJ(i + 1, j, k) - UJ(i, j, k)) * dxiv);
        A3(s22, i, j, k) = jcbiv2 * ((VJ(i, j + 1, k) - VJ(i, j, k)) * dyiv);
        A3(s33, i, j, k) = jcbiv2 * (A3(w, i, j, k + 1) - A3(w, i, j, k)) * dziv;
      }
  }
  for (int k = 2; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1 + iw; i <= ni - ie; i++) {
        const R t3 = K(2.0) / (A3(jcb8w, i - 1, j, k) + A3(jcb8w, i, j, k));
        A3(s13, i, j, k) = ((WJ(i, j, k) - WJ(i - 1, j, k)) * dxiv + (A3(u, i, j, k) - A3(u, i, j, k - 1)) * dziv) * t3;
        A3(s31, i, j, k) = A3(s13, i, j, k);
      }
    for (int j = 1 + js; j <= nj - jn; j++)
      for (int i = 1; i <= ni - 1; i++) {
        const R t3 = K(2.0) / (A3(jcb8w, i, j - 1, k) + A3(jcb8w, i, j, k));
        A3(s23, i, j, k) = ((WJ(i, j, k) - WJ(i, j - 1, k)) * dyiv + (A3(v, i, j, k) - A3(v, i, j, k - 1)) * dziv) * t3;
        A3(s32, i, j, k) = A3(s23, i, j, k);
      }
  }
#undef UJ
#undef VJ
#undef WJ
  or_bc8(c, 0, nk, s12);
  or_bc8(c, 1, nk, s12);
  or_bc8(c, 0, nk, s31);
  or_bc8(c, 1, nk, s32);
  or_bcten(c, s31);
  or_bcten(c, s32);
}

void or_defomssq(const or_cfg *c, const R *s11, const R *s22, const R *s33, const R *s12, const R *s31,
                 const R *s32, R *ssq)
{
  DECL_DIMS(c);
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        const R a12 = (A3(s12, i, j, k) + A3(s12, i + 1, j + 1, k)) + (A3(s12, i + 1, j, k) + A3(s12, i, j + 1, k));
        const R a31 = (A3(s31, i, j, k) + A3(s31, i + 1, j, k + 1)) + (A3(s31, i, j, k + 1) + A3(s31, i + 1, j, k));
        const R a32 = (A3(s32, i, j, k) + A3(s32, i, j + 1, k + 1)) + (A3(s32, i, j + 1, k) + A3(s32, i, j, k + 1));
        A3(ssq, i, j, k) = K(.5) * (A3(s33, i, j, k) * A3(s33, i, j, k) +
                                    (A3(s11, i, j, k) * A3(s11, i, j, k) + A3(s22, i, j, k) * A3(s22, i, j, k))) +
                           K(.0625) * (a12 * a12 + (a31 * a31 + a32 * a32));
      }
}

/* s_eddyvis, tubopt=2 (Deardorff), sfcopt=0, isoopt=1, mfcopt=0. */
void or_eddyvis(const or_cfg *c, R dtb, const R *jcb, const R *rbr, const R *tke, const R *ssq, const R *nsq8w,
                R *rkh, R *rkv, R *priv)
{
  DECL_DIMS(c);
  const R ds3 = (R)c->dx * (R)c->dy * (R)c->dz;
  const R ckm = K(.1), ckmin = K(1.e-6), ckmax = K(.125) / dtb;
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        const R ln0 = EXP(ONED3 * LOG(ds3 * A3(jcb, i, j, k)));
        const R ln02 = ln0 * ln0;
        const R nsq = A3(nsq8w, i, j, k) + A3(nsq8w, i, j, k + 1);
        R ln;
        if (nsq < K(0.0)) ln = ln0;
        else {
          R m = K(.76) * SQRT(A3(tke, i, j, k) / (nsq + C_EPS));
          if (!(m < ln0)) m = ln0;
          ln = K(.1) * ln0;
          if (!(ln > m)) ln = m;
        }
        A3(priv, i, j, k) = K(1.0) + K(2.0) * ln / ln0;
        R kv = ckm * ln * SQRT(A3(tke, i, j, k));
        if (A3(priv, i, j, k) * nsq < A3(ssq, i, j, k)) {
          const R lo = ckmin * ln02;
          if (lo > kv) kv = lo;
        }
        const R hi = ckmax * ln02;
        A3(rkv, i, j, k) = A3(rbr, i, j, k) * (hi < kv ? hi : kv);
        A3(rkh, i, j, k) = A3(rkv, i, j, k);
      }
}

/* The four in turbdrv.f90's order, with the reference's work arrays kept
 * separate (w1..w3 3-D, w4 2-D). */
void or_turbcoef(const or_cfg *c, R dtb, int moist, const R *jcb, const R *jcb8u, const R *jcb8v, const R *jcb8w,
                 const R *pbr, const R *ptbr, const R *rbr, const R *ppp, const R *ptpp, const R *qvp, const R *qallp,
                 const R *up, const R *vp, const R *wp, const R *tkep, R *t11, R *t22, R *t33, R *t12, R *t13,
                 R *t23, R *t31, R *t32, R *ssq, R *nsq8w, R *rkh, R *rkv, R *priv, R *w1, R *w2, R *w3, R *w4)
{
  or_bruntv(c, moist, jcb8w, pbr, ptbr, ppp, ptpp, qvp, qallp, nsq8w, w1, w2, w3, w4);
  or_defomten(c, jcb, jcb8u, jcb8v, jcb8w, up, vp, wp, t11, t22, t33, t12, t13, t23, t31, t32);
  or_defomssq(c, t11, t22, t33, t12, t31, t32, ssq);
  or_eddyvis(c, dtb, jcb, rbr, tkep, ssq, nsq8w, rkh, rkv, priv);
}

/* s_strsten (strsten.f90:150-210), sfcopt=0: the tensor times the eddy
 * viscosities, in place; then s_bcten on t31 and t32. */
void or_strsten(const or_cfg *c, const R *rkh, const R *rkv, R *t11, R *t22, R *t33, R *t12, R *t13, R *t23,
                R *t31, R *t32)
{
  DECL_DIMS(c);
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        A3(t11, i, j, k) = A3(rkh, i, j, k) * A3(t11, i, j, k);
        A3(t22, i, j, k) = A3(rkh, i, j, k) * A3(t22, i, j, k);
        A3(t33, i, j, k) = A3(rkv, i, j, k) * A3(t33, i, j, k);
      }
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(t12, i, j, k) = K(.25) * A3(t12, i, j, k) *
                           ((A3(rkh, i - 1, j - 1, k) + A3(rkh, i, j, k)) + (A3(rkh, i - 1, j, k) + A3(rkh, i, j - 1, k)));
  }
  for (int k = 2; k <= nk - 1; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(t13, i, j, k) = K(.25) * A3(t13, i, j, k) *
                           ((A3(rkv, i - 1, j, k - 1) + A3(rkv, i, j, k)) + (A3(rkv, i - 1, j, k) + A3(rkv, i, j, k - 1)));
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(t23, i, j, k) = K(.25) * A3(t23, i, j, k) *
                           ((A3(rkv, i, j - 1, k - 1) + A3(rkv, i, j, k)) + (A3(rkv, i, j - 1, k) + A3(rkv, i, j, k - 1)));
  }
  for (int k = 2; k <= nk - 1; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(t31, i, j, k) = K(.25) * A3(t31, i, j, k) *
                           ((A3(rkh, i - 1, j, k - 1) + A3(rkh, i, j, k)) + (A3(rkh, i - 1, j, k) + A3(rkh, i, j, k - 1)));
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(t32, i, j, k) = K(.25) * A3(t32, i, j, k) *
                           ((A3(rkh, i, j - 1, k - 1) + A3(rkh, i, j, k)) + (A3(rkh, i, j - 1, k) + A3(rkh, i, j, k - 1)));
  }
  or_bcten(c, t31);
  or_bcten(c, t32);
}

/* s_turbuvw (turbuvw.f90:288-352, 795-859, 1300-1330), trnopt=0,
 * mfcopt=0, called with idoneopt (=1, so the leapfrog branch): the
 * divergence of the stresses overwrites ufrc, vfrc, wfrc over their
 * interiors.  t11 and t22 are rescaled in place as in the reference; the
 * products the reference keeps in tmp1 are formed where used. */
void or_turbuvw(const or_cfg *c, const R *jcb, const R *jcb8u, const R *jcb8v, R *t11, R *t22, const R *t33,
                const R *t12, const R *t13, const R *t23, const R *t31, const R *t32, R *ufrc, R *vfrc, R *wfrc)
{
  DECL_DIMS(c);
  const R dxiv = K(1.0) / (R)c->dx, dyiv = K(1.0) / (R)c->dy, dziv = K(1.0) / (R)c->dz;
  const R dxiv05 = K(.5) * dxiv, dyiv05 = K(.5) * dyiv;
#define TU(i, j, k) ((A3(jcb8u, i, j - 1, k) + A3(jcb8u, i, j, k)) * A3(t12, i, j, k))
#define TV(i, j, k) ((A3(jcb8v, i - 1, j, k) + A3(jcb8v, i, j, k)) * A3(t12, i, j, k))
#define TW1(i, j, k) ((A3(jcb8u, i, j, k - 1) + A3(jcb8u, i, j, k)) * A3(t31, i, j, k))
#define TW2(i, j, k) ((A3(jcb8v, i, j, k - 1) + A3(jcb8v, i, j, k)) * A3(t32, i, j, k))
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 1; i <= ni - 1; i++) A3(t11, i, j, k) = A3(jcb, i, j, k) * A3(t11, i, j, k);
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(ufrc, i, j, k) = ((A3(t11, i, j, k) - A3(t11, i - 1, j, k)) * dxiv + (TU(i, j + 1, k) - TU(i, j, k)) * dyiv05) +
                            (A3(t13, i, j, k + 1) - A3(t13, i, j, k)) * dziv;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++) A3(t22, i, j, k) = A3(jcb, i, j, k) * A3(t22, i, j, k);
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(vfrc, i, j, k) = ((A3(t22, i, j, k) - A3(t22, i, j - 1, k)) * dyiv + (TV(i + 1, j, k) - TV(i, j, k)) * dxiv05) +
                            (A3(t23, i, j, k + 1) - A3(t23, i, j, k)) * dziv;
  for (int k = 2; k <= nk - 1; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++) A3(t11, i, j, k) = TW1(i, j, k);
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++) A3(t22, i, j, k) = TW2(i, j, k);
  }
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(wfrc, i, j, k) = (A3(t33, i, j, k) - A3(t33, i, j, k - 1)) * dziv +
                            ((A3(t11, i + 1, j, k) - A3(t11, i, j, k)) * dxiv05 + (A3(t22, i, j + 1, k) - A3(t22, i, j, k)) * dyiv05);
#undef TU
#undef TV
#undef TW1
#undef TW2
}

/* s_turbdrv (turbdrv.f90:245-287) for the options above. */
void or_turbdrv(const or_cfg *c, R dtb, int moist, const R *jcb, const R *jcb8u, const R *jcb8v, const R *jcb8w,
                const R *pbr, const R *ptbr, const R *rbr, const R *ppp, const R *ptpp, const R *qvp, const R *qallp,
                const R *up, const R *vp, const R *wp, const R *tkep, R *ufrc, R *vfrc, R *wfrc, R *rkh, R *rkv,
                R *priv, R *ssq, R *nsq8w, R *t13, R *t23, R *t33, R *t11, R *t22, R *t12, R *t31, R *t32, R *w1,
                R *w2, R *w3, R *w4)
{
  or_turbcoef(c, dtb, moist, jcb, jcb8u, jcb8v, jcb8w, pbr, ptbr, rbr, ppp, ptpp, qvp, qallp, up, vp, wp, tkep, t11,
              t22, t33, t12, t13, t23, t31, t32, ssq, nsq8w, rkh, rkv, priv, w1, w2, w3, w4);
  or_strsten(c, rkh, rkv, t11, t22, t33, t12, t13, t23, t31, t32);
  or_turbuvw(c, jcb, jcb8u, jcb8v, t11, t22, t33, t12, t13, t23, t31, t32, ufrc, vfrc, wfrc);
}

/* s_turbflx (turbflx.f90:180-262), trnopt=0, sfcopt=0: the turbulent
 * fluxes of a scalar s at u, v and w points.  The reference's jcbs
 * (s*jcb) is formed where used. */
void or_turbflx(const or_cfg *c, const R *jcb, const R *s, const R *rkh8u, const R *rkh8v, const R *rkv8w, R *h1,
                R *h2, R *h3)
{
  DECL_DIMS(c);
  const R dxiv = K(1.0) / (R)c->dx, dyiv = K(1.0) / (R)c->dy, dziv = K(1.0) / (R)c->dz;
  const R dxv05 = K(.5) * dxiv, dyv05 = K(.5) * dyiv;
#define JS(i, j, k) (A3(s, i, j, k) * A3(jcb, i, j, k))
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(h1, i, j, k) = A3(rkh8u, i, j, k) * (JS(i, j, k) - JS(i - 1, j, k)) * dxv05;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(h2, i, j, k) = A3(rkh8v, i, j, k) * (JS(i, j, k) - JS(i, j - 1, k)) * dyv05;
  }
#undef JS
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(h3, i, j, k) = A3(rkv8w, i, j, k) * (A3(s, i, j, k) - A3(s, i, j, k - 1)) * dziv;
}

/* s_turbs (turbs.f90:189-212), trnopt=0, mfcopt=0: the divergence of the
 * turbulent fluxes h1, h2, h3 added to sfrc.  The reference's tmp1/tmp2
 * (jcb8u*h1, jcb8v*h2) are formed where used. */
void or_turbs(const or_cfg *c, const R *jcb8u, const R *jcb8v, const R *h1, const R *h2, const R *h3, R *sfrc)
{
  DECL_DIMS(c);
  const R dxiv = K(1.0) / (R)c->dx, dyiv = K(1.0) / (R)c->dy, dziv = K(1.0) / (R)c->dz;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(sfrc, i, j, k) = A3(sfrc, i, j, k) +
                            ((A3(h3, i, j, k + 1) - A3(h3, i, j, k)) * dziv +
                             ((A3(jcb8u, i + 1, j, k) * A3(h1, i + 1, j, k) - A3(jcb8u, i, j, k) * A3(h1, i, j, k)) * dxiv +
                              (A3(jcb8v, i, j + 1, k) * A3(h2, i, j + 1, k) - A3(jcb8v, i, j, k) * A3(h2, i, j, k)) * dyiv));
}

/* s_kh8uv (kh8uv.f90:214-232), mfcopt=0: twice the mean eddy viscosity at
 * u and v points (the halving is folded into the flux coefficients). */
void or_kh8uv(const or_cfg *c, const R *rkh, R *rkh8u, R *rkh8v)
{
  DECL_DIMS(c);
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 1; i++) A3(rkh8u, i, j, k) = A3(rkh, i - 1, j, k) + A3(rkh, i, j, k);
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(rkh8v, i, j, k) = A3(rkh, i, j - 1, k) + A3(rkh, i, j, k);
  }
}

/* s_forcetke (forcetke.f90:225-351): the forcing of the TKE equation for
 * tubopt=2, isoopt=1, trnopt=0, mfcopt=0, no smoothing (smtopt=0) and no
 * lateral sponge, with the vertical sponge s_vsps0 when the tenth letter
 * of vspvar is 'o' and vspopt >= 1.  The calls in the reference's order:
 *   s_advs     advection of the present TKE      (advs.f90:205-445)
 *   s_sheartke +jcb*rkv*ssq                       (sheartke.f90:93-101)
 *   s_disptke  -(.37 priv-.18) rst e^1.5 / l      (disptke.f90:180-195)
 *   s_eddyvisj rkh, rkv divided by jcb            (eddyvisj.f90:103-112)
 *   s_tkeflx   the fluxes of jcb e and e          (tkeflx.f90:181-330)
 *   s_turbtke  their divergence                   (turbtke.f90:191-213)
 *   s_vsps0    -.5 (rbct_k + rbct_k+1) rst e      (vsps0.f90:114-127)
 *   s_eddydif  rkh, rkv8w, rkv8s for the scalars  (eddydif.f90:266-330)
 *   s_buoytke  -(jcb8w N^2 (rkv8s_k-1 + rkv8s_k))  (buoytke.f90:96-115)
 * rkh and rkv are updated in place as the reference updates them; rkv8s
 * is an output; w1..w4 are work arrays (the reference's h3, tmp1..tmp3,
 * and ssq, which it also uses as work and which is read here only). */
void or_forcetke(const or_cfg *c, const R *jcb, const R *jcb8u, const R *jcb8v, const R *jcb8w, const R *rst,
                 const R *rstxu, const R *rstxv, const R *rstxwc, const R *tke, const R *tkep, const R *rbct,
                 const R *priv, const R *nsq8w, const R *ssq, R *rkh, R *rkv, R *tkefrc, R *rkv8s, R *w1, R *w2,
                 R *w3, R *w4)
{
  DECL_DIMS(c);
  or_advs(c, rstxu, rstxv, rstxwc, tke, tkefrc, w1, w2, w3, w4);
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tkefrc, i, j, k) = A3(tkefrc, i, j, k) + A3(jcb, i, j, k) * A3(rkv, i, j, k) * A3(ssq, i, j, k);
  const R ds308 = K(.125) * (R)c->dx * (R)c->dy * (R)c->dz;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++) {
        const R pv = A3(priv, i, j, k), e = A3(tkep, i, j, k);
        const R ln = (pv - K(1.0)) * EXP(ONED3 * LOG(ds308 * A3(jcb, i, j, k))) + C_EPS;
        A3(tkefrc, i, j, k) = A3(tkefrc, i, j, k) - (K(.37) * pv - K(.18)) * A3(rst, i, j, k) * e * SQRT(e) / ln;
      }
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        const R jcbiv = K(1.0) / A3(jcb, i, j, k);
        A3(rkh, i, j, k) = jcbiv * A3(rkh, i, j, k);
        A3(rkv, i, j, k) = jcbiv * A3(rkv, i, j, k);
      }
  /* s_tkeflx: h1, h2, h3 in w1, w2, w3; jcb*e formed where used */
  const R dxiv = K(1.0) / (R)c->dx, dyiv = K(1.0) / (R)c->dy, dziv = K(1.0) / (R)c->dz;
  const R dxv5 = K(.5) * dxiv, dyv5 = K(.5) * dyiv, dzv5 = K(.5) * dziv;
#define JT(i, j, k) (A3(tkep, i, j, k) * A3(jcb, i, j, k))
  for (int k = 1; k <= nk - 1; k++) {
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 1; i++)
        A3(w1, i, j, k) = (A3(rkh, i - 1, j, k) + A3(rkh, i, j, k)) * (JT(i, j, k) - JT(i - 1, j, k)) * dxv5;
    for (int j = 2; j <= nj - 1; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(w2, i, j, k) = (A3(rkh, i, j - 1, k) + A3(rkh, i, j, k)) * (JT(i, j, k) - JT(i, j - 1, k)) * dyv5;
  }
#undef JT
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(w3, i, j, k) = (A3(rkv, i, j, k - 1) + A3(rkv, i, j, k)) * (A3(tkep, i, j, k) - A3(tkep, i, j, k - 1)) * dzv5;
  /* s_turbtke: jcb8u*h1 and jcb8v*h2 formed where used */
  const R dxiv2 = K(2.0) * dxiv, dyiv2 = K(2.0) * dyiv, dziv2 = K(2.0) * dziv;
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tkefrc, i, j, k) =
            A3(tkefrc, i, j, k) +
            ((A3(w3, i, j, k + 1) - A3(w3, i, j, k)) * dziv2 +
             ((A3(jcb8u, i + 1, j, k) * A3(w1, i + 1, j, k) - A3(jcb8u, i, j, k) * A3(w1, i, j, k)) * dxiv2 +
              (A3(jcb8v, i, j + 1, k) * A3(w2, i, j + 1, k) - A3(jcb8v, i, j, k) * A3(w2, i, j, k)) * dyiv2));
  if (c->vspopt >= 1 && c->vspvar[9] == 'o') or_vsps(c, rst, tkep, NULL, rbct, tkefrc);
  /* s_eddydif, tubopt>=2, isoopt=1 */
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        A3(rkv8s, i, j, k) = A3(priv, i, j, k) * A3(rkv, i, j, k);
        A3(rkh, i, j, k) = A3(rkv8s, i, j, k);
      }
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(rkv, i, j, k) = K(.5) * (A3(rkv8s, i, j, k - 1) + A3(rkv8s, i, j, k));
  for (int k = 1; k <= nk - 1; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) A3(rkv8s, i, j, k) = K(.5) * A3(jcb, i, j, k) * A3(rkv8s, i, j, k);
  /* s_buoytke: its tmp1 in w4 */
  for (int k = 2; k <= nk - 1; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(w4, i, j, k) = A3(jcb8w, i, j, k) * A3(nsq8w, i, j, k) * (A3(rkv8s, i, j, k - 1) + A3(rkv8s, i, j, k));
  for (int k = 2; k <= nk - 2; k++)
    for (int j = 2; j <= nj - 2; j++)
      for (int i = 2; i <= ni - 2; i++)
        A3(tkefrc, i, j, k) = A3(tkefrc, i, j, k) - (A3(w4, i, j, k) + A3(w4, i, j, k + 1));
}
