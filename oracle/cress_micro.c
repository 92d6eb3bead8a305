/* Cold-rain bulk microphysics (cphopt=2, haiopt=0) of the oracle -- TEST
 * INFRASTRUCTURE ONLY; #included by cress_oracle.c.
 *
 * Every routine restates the reference routine cited at its definition over
 * the reference's own loop range k=1..nk-1, j=1..nj-1, i=1..ni-1, with the
 * reference's operation order (Fortran left-to-right, explicit parentheses)
 * so that the FP32 build reproduces the reference's FP32 results; exp/log/
 * sqrt are the C library's, as the reference's FP32 build on the same host
 * calls them.  Options fixed by the configuration: flqcqi_opt=1,
 * evapor_opt=1, nuc2nd_opt=1 (SrcDevelop/0rig/UserModAdd/temparam.f90),
 * advopt<=3 (leapfrog: microphysics dt = 2*dtb, cloudphy.f90:281-285).
 */

/* comphy.f90 / commath.f90 constants, in REAL */
#define C_G K(9.8)
#define C_CP K(1004.0)
#define C_CW K(4.218e3)
#define C_CI K(2.106e3)
#define C_RD K(287.0)
#define C_RV K(461.0)
#define C_EPSVA K(0.622)
#define C_ES0 K(610.78)
#define C_R0 K(1.225)
#define C_P0 K(1.e5)
#define C_T0 K(273.16)
#define C_TLOW K(233.16)
#define C_T0CEL K(0.0)
#define C_T20 K(293.16)
#define C_T23 K(296.16)
#define C_P20 K(101325.0)
#define C_LMB20 K(6.6e-8)
#define C_DIASL K(6.e-7)
#define C_BOLTZ K(1.38066e-23)
#define C_SC K(0.6)
#define C_LV0 K(2.50078e6)
#define C_LF0 K(3.34e5)
#define C_KP0 K(10.79)
#define C_KPA K(2.4e-1)
#define C_NU0 K(1.328e-5)
#define C_DV0 K(2.23e-5)
#define C_XL K(0.25)
#define C_NR0 K(8.e6)
#define C_NS0 K(1.8e6)
#define C_NG0 K(1.1e6)
#define C_AUC K(2.98e7)
#define C_AUR K(842.0)
#define C_AUI K(770.0)
#define C_AUS K(27.0)
#define C_AUG K(132.0)
#define C_BUC K(2.0)
#define C_BUR K(0.8)
#define C_BUI K(1.0)
#define C_BUS K(0.5)
#define C_BUG K(0.64)
#define C_GUC K(1.0)
#define C_GUR K(0.5)
#define C_GUI K(0.33)
#define C_GUS K(0.5)
#define C_GUG K(0.5)
#define C_RHOW K(1.e3)
#define C_RHOI K(2.e2)
#define C_RHOS K(1.e2)
#define C_RHOG K(4.e2)
#define C_ERS K(1.0)
#define C_ERG K(1.0)
#define C_EIR K(1.0)
#define C_EIS K(0.1)
#define C_EIG K(0.1)
#define C_ESR K(1.0)
#define C_ESG K(0.001)
#define C_ECC K(0.55)
#define C_EII K(0.1)
#define C_ESS K(0.1)
#define C_MR0 K(2.7e-10)
#define C_MI0 K(1.e-12)
#define C_MS0 K(1.76e-10)
#define C_MG0 K(7.e-10)
#define C_MRMAX K(2.68e-4)
#define C_MIMAX K(1.4e-10)
#define C_MSMAX K(6.e-4)
#define C_MGMAX K(8.e-3)
#define C_NCLCST K(1.e8)
#define C_DIAQCM K(2.e-5)
#define C_EPS K(1.e-20)
#define C_CC K(3.141592)
#define C_LN10 K(2.30259)
#define C_GF4 K(6.0)
#define C_GF7 K(720.0)
#define C_GFBUS K(1.77245)
#define C_GF1BUC K(2.0)
#define C_GF1BUR K(0.931384)
#define C_GF1BUS K(0.886227)
#define C_GF1BUG K(0.898642)
#define C_GF3BUR K(4.69417)
#define C_GF3BUS K(3.32335)
#define C_GF3BUG K(3.89076)
#define C_GF4BUC K(120.0)
#define C_GF4BUR K(17.8379)
#define C_GF4BUS K(11.6314)
#define C_GF4BUG K(14.1624)
#define C_GF5BUR K(1.82736)
#define C_GF5BUS K(1.60836)
#define C_GF5BUG K(1.70506)
#define C_HFBUS K(1610.0)
#define ONED3 ((R)1 / (R)3)
#define ONED6 ((R)1 / (R)6)
#define ONED9 ((R)1 / (R)9)
#define SEVND3 ((R)7 / (R)3)
/* temparam.f90 (UserModAdd overlay) */
#define C_UCQCST K(0.1)
#define C_UCNCST K(0.1)
#define C_UIQCST K(0.0001)
#define C_UINCST K(0.0001)

static R rmin(R a, R b) { return a < b ? a : b; }
static R rmax(R a, R b) { return a > b ? a : b; }

/* Koenig (1971) depositional growth parameters, setref.f90:473-600:
 * a table of published constants, kept as the reference lists it. */
static R g_ckoe[41], g_pkoe[41];
static int g_koe_ready = 0;
static void koenig_table(void)
{
  const R c[41] = {
      K(0.0), K(.7939e-10), K(.7841e-9), K(.3369e-8), K(.4336e-8), K(.5285e-8), K(.3728e-8),
      K(.1852e-8), K(.2991e-9), K(.4248e-9), K(.7434e-9), K(.1812e-8), K(.4394e-8), K(.9145e-8),
      K(.1725e-7), K(.3348e-7), K(.1725e-7), K(.9175e-8), K(.4412e-8), K(.2252e-8), K(.9115e-9),
      K(.4876e-9), K(.3473e-9), K(.4758e-9), K(.6306e-9), K(.8573e-9), K(.7868e-9), K(.7192e-9),
      K(.6513e-9), K(.5956e-9), K(.5333e-9), K(.4834e-9), K(.4335e-9), K(.3836e-9), K(.3337e-9),
      K(.2838e-9), K(.2339e-9), K(.1840e-9), K(.1341e-9), K(.8420e-10), K(.3430e-10)};
  const R p[41] = {
      K(0.0), K(.4006), K(.4831), K(.5320), K(.5307), K(.5319), K(.5249), K(.4888), K(.3894),
      K(.4047), K(.4318), K(.4771), K(.5183), K(.5463), K(.5651), K(.5813), K(.5655), K(.5478),
      K(.5203), K(.4906), K(.4447), K(.4126), K(.3960), K(.4149), K(.4320), K(.4506), K(.4483),
      K(.4460), K(.4433), K(.4413), K(.4382), K(.4361), K(.4340), K(.4319), K(.4298), K(.4277),
      K(.4256), K(.4235), K(.4214), K(.4193), K(.4172)};
  if (g_koe_ready) return;
  for (int t = 0; t <= 40; t++) {
    g_pkoe[t] = p[t];
    g_ckoe[t] = c[t] * EXP(K(3.0) * p[t] * C_LN10);
  }
  g_koe_ready = 1;
}

/* Work arrays of one cloudphy call: 3-D fields, one per reference array. */
enum {
  MW_RBV, MW_PI, MW_P, MW_NCC, MW_NCR, MW_NCS, MW_NCG,
  MW_UCQ, MW_URQ, MW_UIQ, MW_USQ, MW_UGQ, MW_UCN, MW_URN, MW_UIN, MW_USN, MW_UGN,
  MW_T, MW_TCEL, MW_QVSST0, MW_QVSW, MW_QVSI, MW_LV, MW_LS, MW_LF, MW_KP, MW_MU, MW_DV, MW_MI,
  MW_DIAQC, MW_DIAQR, MW_DIAQI, MW_DIAQS, MW_DIAQG, MW_VNTR, MW_VNTS, MW_VNTG,
  MW_NUVI, MW_NUCI, MW_CLCR, MW_CLCS, MW_CLCG, MW_CLRI, MW_CLRS, MW_CLRG, MW_CLIR, MW_CLIS, MW_CLIG,
  MW_CLSR, MW_CLSG, MW_CLRSG, MW_AGIN, MW_VDVR, MW_VDVI, MW_VDVS, MW_VDVG, MW_CNCR, MW_CNIS, MW_CNSG,
  MW_SPSI, MW_SPGI, MW_MLIC, MW_MLSR, MW_MLGR, MW_FRRG, MW_SHSR, MW_SHGR, MW_PGWET, MW_ECS, MW_FLX,
  MW_NW
};

#define MLOOP_SERIAL                            \
  for (int k = 1; k <= nk - 1; k++)             \
    for (int j = 1; j <= nj - 1; j++)           \
      for (int i = 1; i <= ni - 1; i++)
/* pointwise sweeps: levels across OpenMP threads (same per-cell result) */
#define MLOOP                                   \
  OMP_FOR                                       \
  for (int k = 1; k <= nk - 1; k++)             \
    for (int j = 1; j <= nj - 1; j++)           \
      for (int i = 1; i <= ni - 1; i++)
#define PIX ((size_t)i + NX * ((size_t)j + NY * ((size_t)k - 1)))

/* s_getvdens, getvdens.f90:100-106 */
void or_getvdens(const or_cfg *c, const R *rbr, R *rbv)
{
  DECL_DIMS(c);
  MLOOP { size_t p = PIX; rbv[p] = K(1.0) / rbr[p]; }
}

/* s_getexner, getexner.f90:113-131 */
void or_getexner(const or_cfg *c, const R *pbr, const R *pp, R *pi, R *p)
{
  DECL_DIMS(c);
  const R rddvcp = C_RD / C_CP, p0iv = K(1.0) / C_P0;
  MLOOP {
    size_t q = PIX;
    p[q] = pbr[q] + pp[q];
    pi[q] = EXP(rddvcp * LOG(p0iv * p[q]));
  }
}

/* s_diagnw, diagnw.f90:120-156: ncc = nclcst/rho, ncr from qr */
void or_diagnw(const or_cfg *c, const R *rbr, const R *qc, const R *qr, R *ncc, R *ncr)
{
  DECL_DIMS(c);
  (void)qc;
  const R mrmiv2 = K(1.e-2) / C_MRMAX, mr0iv2 = K(1.e2) / C_MR0;
  const R cdiaqr = C_NR0 * C_NR0 * C_NR0 / (C_CC * C_RHOW);
  MLOOP {
    size_t p = PIX;
    R rbv = K(1.0) / rbr[p];
    ncc[p] = C_NCLCST * rbv;
    R n = SQRT(SQRT(cdiaqr * rbr[p] * qr[p])) * rbv;
    ncr[p] = rmin(rmax(n, mrmiv2 * qr[p]), mr0iv2 * qr[p]);
  }
}

/* s_diagnsg, diagnsg.f90:147-194 (haiopt=0) */
void or_diagnsg(const or_cfg *c, const R *rbr, const R *rbv, const R *qs, const R *qg, R *ncs, R *ncg)
{
  DECL_DIMS(c);
  const R msmiv2 = K(1.e-2) / C_MSMAX, ms0iv2 = K(1.e2) / C_MS0;
  const R mgmiv2 = K(1.e-2) / C_MGMAX, mg0iv2 = K(1.e2) / C_MG0;
  const R cdiaqs = C_NS0 * C_NS0 * C_NS0 / (C_CC * C_RHOS);
  const R cdiaqg = C_NG0 * C_NG0 * C_NG0 / (C_CC * C_RHOG);
  MLOOP {
    size_t p = PIX;
    R a = SQRT(SQRT(cdiaqs * rbr[p] * qs[p])) * rbv[p];
    ncs[p] = rmin(rmax(a, msmiv2 * qs[p]), ms0iv2 * qs[p]);
    R b = SQRT(SQRT(cdiaqg * rbr[p] * qg[p])) * rbv[p];
    ncg[p] = rmin(rmax(b, mgmiv2 * qg[p]), mg0iv2 * qg[p]);
  }
}

/* s_adjstnci, adjstnci.f90:110-126 */
void or_adjstnci(const or_cfg *c, const R *qi, R *nci)
{
  DECL_DIMS(c);
  const R mimiv5 = K(0.5) / C_MIMAX, mi0iv2 = K(1.e2) / C_MI0;
  MLOOP {
    size_t p = PIX;
    nci[p] = rmin(rmax(nci[p], mimiv5 * qi[p]), mi0iv2 * qi[p]);
  }
}

/* s_termblk, termblk.f90:267-493 (flqcqi_opt=1, haiopt=0): terminal
 * velocities of mass (u?q) and number (u?n) from the past level. */
void or_termblk(const or_cfg *c, const R *rbv, const R *qc, const R *qr, const R *qi, const R *qs,
                const R *qg, const R *ncr, const R *ncs, const R *ncg, R *ucq, R *urq, R *uiq,
                R *usq, R *ugq, R *ucn, R *urn, R *uin, R *usn, R *ugn)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R bur3 = ONED3 * C_BUR, bus3 = ONED3 * C_BUS, bug3 = ONED3 * C_BUG;
  const R curq = ONED6 * C_AUR * C_GF4BUR, cusq = ONED6 * C_AUS * C_GF4BUS, cugq = ONED6 * C_AUG * C_GF4BUG;
  const R curn = K(6.0) * C_GF1BUR / C_GF4BUR, cusn = K(6.0) * C_GF1BUS / C_GF4BUS;
  const R cugn = K(6.0) * C_GF1BUG / C_GF4BUG;
  const R cdiaqr = K(1.0) / (C_CC * C_RHOW), cdiaqs = K(1.0) / (C_CC * C_RHOS);
  const R cdiaqg = K(1.0) / (C_CC * C_RHOG);
  MLOOP {
    size_t p = PIX;
    if (qc[p] > thresq) { ucq[p] = C_UCQCST; ucn[p] = C_UCNCST; }
    else { ucq[p] = K(0.0); ucn[p] = K(0.0); }
    if (qi[p] > thresq) { uiq[p] = C_UIQCST; uin[p] = C_UINCST; }
    else { uiq[p] = K(0.0); uin[p] = K(0.0); }
  }
  MLOOP {
    size_t p = PIX;
    R lnr0r = LOG(C_R0 * rbv[p]);
    if (qr[p] > thresq) {
      urq[p] = curq * EXP(C_GUR * lnr0r) * EXP(bur3 * LOG(cdiaqr * qr[p] / ncr[p]));
      urn[p] = curn * urq[p];
    } else { urq[p] = K(0.0); urn[p] = K(0.0); }
    if (qs[p] > thresq) {
      usq[p] = cusq * EXP(C_GUS * lnr0r) * EXP(bus3 * LOG(cdiaqs * qs[p] / ncs[p]));
      usn[p] = cusn * usq[p];
    } else { usq[p] = K(0.0); usn[p] = K(0.0); }
    if (qg[p] > thresq) {
      ugq[p] = cugq * EXP(C_GUG * lnr0r) * EXP(bug3 * LOG(cdiaqg * qg[p] / ncg[p]));
      ugn[p] = cugn * ugq[p];
    } else { ugq[p] = K(0.0); ugn[p] = K(0.0); }
  }
}

/* Arrays of one srcblk call (the reference's argument list, srcblk.f90). */
typedef struct {
  const R *ptbr, *rbr, *rbv, *pi, *p, *ptpp, *qvp, *qcp, *qrp, *qip, *qsp, *qgp;
  const R *nccp, *ncrp, *ncip, *ncsp, *ncgp, *urq, *usq, *ugq;
  R *ptpf, *qvf, *qcf, *qrf, *qif, *qsf, *qgf, *ncif;
  R *w[MW_NW];   /* work arrays indexed by M_* */
} or_src;

#define W(name) s->w[MW_##name]

/* s_setblk, setblk.f90:269-943 (cphopt<=3, 3-D branch) */
static void or_setblk(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R cwmci = C_CW - C_CI, epses0 = C_EPSVA * C_ES0, t23iv = K(1.0) / C_T23;
  R cdv = K(1.0) / C_T0;
  const R cnu = EXP(K(.57013) * LOG(C_NU0)) * cdv;
  cdv = EXP(K(.55249) * LOG(C_DV0)) * cdv;
  const R pdiaqr = K(.5) * (K(5.0) + C_BUR) - K(1.0);
  const R pdiaqs = K(.5) * (K(5.0) + C_BUS) - K(1.0);
  const R pdiaqg = K(.5) * (K(5.0) + C_BUG) - K(1.0);
  const R ccrw6 = K(6.0) / (C_CC * C_RHOW), ccri6 = K(6.0) / (C_CC * C_RHOI);
  const R cdiaqr = K(1.0) / (C_CC * C_RHOW), cdiaqs = K(1.0) / (C_CC * C_RHOS);
  const R cdiaqg = K(1.0) / (C_CC * C_RHOG);
  const R cvntr = K(.31) * SQRT(C_AUR) * EXP(ONED3 * LOG(C_SC)) * C_GF5BUR;
  const R cvnts = K(.31) * SQRT(C_AUS) * EXP(ONED3 * LOG(C_SC)) * C_GF5BUS;
  const R cvntg = K(.31) * SQRT(C_AUG) * EXP(ONED3 * LOG(C_SC)) * C_GF5BUG;
  OMP_FOR
  for (int k = 1; k <= nk - 1; k++) {
    R *nu = W(FLX) + NX * NY * (size_t)(k - 1);   /* nu(i,j): the level's 2-D scratch */
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        size_t p = PIX;
        R t = (s->ptbr[p] + s->ptpp[p]) * s->pi[p];
        W(T)[p] = t;
        R tcel = t - C_T0;
        W(TCEL)[p] = tcel;
        R esw = C_ES0 * EXP(K(17.269) * tcel / (t - K(35.86)));
        R esi = C_ES0 * EXP(K(21.875) * tcel / (t - K(7.66)));
        W(QVSST0)[p] = s->qvp[p] - epses0 / (s->p[p] - C_ES0);
        W(QVSW)[p] = C_EPSVA * esw / (s->p[p] - esw);
        W(QVSI)[p] = C_EPSVA * esi / (s->p[p] - esi);
        R lv = C_LV0 * EXP((K(.167) + K(3.67e-4) * t) * LOG(C_T0 / t));
        W(LV)[p] = lv;
        R lf = C_LF0 + cwmci * tcel;
        W(LF)[p] = lf;
        W(LS)[p] = lv + lf;
        R p20dvp = C_P20 / s->p[p];
        size_t q2 = (size_t)i + NX * (size_t)j;
        nu[q2] = p20dvp * EXP(K(1.754) * LOG(cnu * t));
        W(KP)[p] = (C_KP0 / (K(120.0) + t)) * EXP(K(1.5) * LOG(t23iv * t));
        W(MU)[p] = s->rbr[p] * nu[q2];
        W(DV)[p] = p20dvp * EXP(K(1.81) * LOG(cdv * t));
      }
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        size_t p = PIX;
        size_t q2 = (size_t)i + NX * (size_t)j;
        R cvnt = SQRT(SQRT(C_R0 * s->rbv[p]) / nu[q2]);
        if (s->qcp[p] > thresq) W(DIAQC)[p] = EXP(ONED3 * LOG(ccrw6 * s->qcp[p] / s->nccp[p]));
        else W(DIAQC)[p] = K(0.0);
        if (s->qrp[p] > thresq) {
          R d = EXP(ONED3 * LOG(cdiaqr * s->qrp[p] / s->ncrp[p]));
          W(DIAQR)[p] = d;
          W(VNTR)[p] = s->rbr[p] * s->ncrp[p] * (K(.78) * d + cvnt * cvntr * EXP(pdiaqr * LOG(d)));
        } else { W(DIAQR)[p] = K(0.0); W(VNTR)[p] = K(0.0); }
        if (s->qip[p] > thresq) {
          R mi = s->qip[p] / s->ncip[p];
          W(MI)[p] = mi;
          W(DIAQI)[p] = EXP(ONED3 * LOG(ccri6 * mi));
        } else { W(MI)[p] = K(0.0); W(DIAQI)[p] = K(0.0); }
        if (s->qsp[p] > thresq) {
          R d = EXP(ONED3 * LOG(cdiaqs * s->qsp[p] / s->ncsp[p]));
          W(DIAQS)[p] = d;
          W(VNTS)[p] = s->rbr[p] * s->ncsp[p] * (K(.78) * d + cvnt * cvnts * EXP(pdiaqs * LOG(d)));
        } else { W(DIAQS)[p] = K(0.0); W(VNTS)[p] = K(0.0); }
        if (s->qgp[p] > thresq) {
          R d = EXP(ONED3 * LOG(cdiaqg * s->qgp[p] / s->ncgp[p]));
          W(DIAQG)[p] = d;
          W(VNTG)[p] = s->rbr[p] * s->ncgp[p] * (K(.78) * d + cvnt * cvntg * EXP(pdiaqg * LOG(d)));
        } else { W(DIAQG)[p] = K(0.0); W(VNTG)[p] = K(0.0); }
      }
  }
}

/* s_nuc1stv, nuc1stv.f90:179-216: deposition/sorption nucleation of ice */
static void or_nuc1stv(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  MLOOP {
    size_t p = PIX;
    R qv = s->qvp[p], qvsi = W(QVSI)[p], t = W(T)[p];
    if (qv > thresq && qv > qvsi && t > C_TLOW && t < C_T0) {
      R qvssi = qv - qvsi;
      R a = K(15.25) * qv / qvsi - K(10.08);
      if (a < K(40.0))
        W(NUVI)[p] = rmax(K(0.0), rmin(C_MI0 * EXP(a) * s->rbv[p] - s->qip[p], qvssi));
      else
        W(NUVI)[p] = qvssi;
    } else {
      W(NUVI)[p] = K(0.0);
    }
  }
}

/* s_nuc1stc, nuc1stc.f90:172-401: freezing/contact nucleation of cloud water */
static void or_nuc1stc(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R rwdt2 = (K(100.0) / C_RHOW) * dtb;
  const R cc45 = K(4.e5) * C_CC;
  const R cknd = K(2.0) * C_LMB20 * C_P20 / (C_T20 * C_DIASL);
  const R cdar = C_BOLTZ / (K(3.0) * C_CC * C_DIASL);
  const R kpa25 = K(2.5) * C_KPA, kpa50 = K(5.0) * C_KPA;
  MLOOP {
    size_t p = PIX;
    R qc = s->qcp[p], t = W(T)[p];
    if (qc > thresq) {
      R nufci, nucci, nuhci;
      if (t > C_TLOW && t < C_T0)
        nufci = rmin(rwdt2 * (EXP(K(-.66) * W(TCEL)[p]) - K(1.0)) * qc * qc / s->nccp[p], qc);
      else
        nufci = K(0.0);
      R tc = t;
      if (tc > C_TLOW && tc < K(270.16)) {
        R piv = K(1.0) / s->p[p];
        R knd = cknd * piv * t;
        R dar = cdar * tc * (K(1.0) + knd) / W(MU)[p];
        R f1 = cc45 * EXP(K(1.3) * LOG(K(270.16) - tc)) * W(DIAQC)[p];
        if (tc < t) {
          R f2 = C_KPA * piv * (t - tc);
          R ft = (W(KP)[p] + kpa25 * knd) * (K(.4) + K(.58) * knd + K(.16) * EXP(K(-1.0) / knd)) /
                 ((K(1.0) + K(3.0) * knd) * (K(2.0) * W(KP)[p] + kpa50 * knd + C_KPA));
          nucci = f1 * f2 * (C_RV * t / W(LV)[p] + ft);
        } else {
          nucci = K(0.0);
        }
        nucci = (nucci + f1 * dar) * qc * dtb;
      } else {
        nucci = K(0.0);
      }
      nuhci = t <= C_TLOW ? qc : K(0.0);
      W(NUCI)[p] = nufci + nucci + nuhci;
    } else {
      W(NUCI)[p] = K(0.0);
    }
  }
}

/* s_collect, collect.f90:297-321 constants, 1164-1534 (cphopt=2, 3-D) */
static void or_collect(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R rwdv9 = ONED9 * C_RHOW;
  const R bur2 = K(2.0) + C_BUR, bus2 = K(2.0) + C_BUS, bug2 = K(2.0) + C_BUG;
  const R cclcr = K(.25) * C_AUR * C_GF3BUR * C_CC * dtb;
  const R cclcs = K(.25) * C_AUS * C_GF3BUS * C_CC * dtb;
  const R cclcg = K(.25) * C_AUG * C_GF3BUG * C_CC * dtb;
  const R cclri = K(.25) * C_AUR * C_GF3BUR * C_EIR * C_CC * dtb;
  const R cclrs = C_ERS * C_CC * C_CC * C_RHOW * dtb;
  const R cclrg = C_ERG * C_CC * C_CC * C_RHOW * dtb;
  const R cclis = K(.25) * C_AUS * C_GF3BUS * C_EIS * C_CC * dtb;
  const R cclig = K(.25) * C_AUG * C_GF3BUG * C_EIG * C_CC * dtb;
  const R cclsr = C_ESR * C_CC * C_CC * C_RHOS * dtb;
  const R cclsg = C_ESG * C_CC * C_CC * C_RHOS * dtb;
  const R cclsg3 = C_CC * C_CC * C_RHOS * dtb;
  MLOOP {
    size_t p = PIX;
    const R qc = s->qcp[p], qr = s->qrp[p], qi = s->qip[p], qs = s->qsp[p], qg = s->qgp[p];
    const R rbr = s->rbr[p], t = W(T)[p], mu = W(MU)[p];
    const R dr = W(DIAQR)[p], ds = W(DIAQS)[p], dg = W(DIAQG)[p];
    const R urq = s->urq[p], usq = s->usq[p], ugq = s->ugq[p];
    const R ncr = s->ncrp[p], ncs = s->ncsp[p], ncg = s->ncgp[p];
    R r0rsq = SQRT(C_R0 * rbr);
    R diaqr2, diaqr3, qr2b, diaqs2, diaqs3, qs2b, diaqg2, qg2b, a, b, sink;
    if (qr > thresq) { diaqr2 = dr * dr; diaqr3 = dr * diaqr2; qr2b = r0rsq * EXP(bur2 * LOG(dr)); }
    else { diaqr2 = K(0.0); diaqr3 = K(0.0); qr2b = K(0.0); }
    if (qs > thresq) { diaqs2 = ds * ds; diaqs3 = ds * diaqs2; qs2b = r0rsq * EXP(bus2 * LOG(ds)); }
    else { diaqs2 = K(0.0); diaqs3 = K(0.0); qs2b = K(0.0); }
    if (qg > thresq) { diaqg2 = dg * dg; qg2b = r0rsq * EXP(bug2 * LOG(dg)); }
    else { diaqg2 = K(0.0); qg2b = K(0.0); }
    R clcr = 0, clcs = 0, clcg = 0, clri = 0, clrs = 0, clrg = 0, clir = 0, clsr = 0, ecs = 0;
    R clis, clig, clsg;
    if (t > C_TLOW) {
      if (qc > thresq) {
        R cstk = rwdv9 * W(DIAQC)[p] * W(DIAQC)[p];
        if (qr > thresq) {
          a = cstk * urq / (dr * mu);
          b = a + K(.5);
          a = a * a / (b * b);
          clcr = cclcr * a * qr2b * ncr * qc;
        } else clcr = K(0.0);
        if (qs > thresq) {
          a = cstk * usq / (ds * mu);
          b = a + K(.5);
          ecs = a * a / (b * b);
          clcs = cclcs * qs2b * ecs * ncs * qc;
        } else { clcs = K(0.0); ecs = K(0.0); }
        if (qg > thresq) {
          a = cstk * ugq / (dg * mu);
          b = a + K(.5);
          a = a * a / (b * b);
          clcg = cclcg * a * qg2b * ncg * qc;
        } else clcg = K(0.0);
        sink = clcr + clcs + clcg;
        if (qc < sink) { a = qc / sink; clcr = clcr * a; clcs = clcs * a; clcg = clcg * a; }
      } else { clcr = clcs = clcg = ecs = K(0.0); }
      if (qr > thresq) {
        if (qi > thresq && t < C_T0) { clri = cclri * qr2b * ncr * qi; clir = clri; }
        else { clri = K(0.0); clir = K(0.0); }
        if (qs > thresq) {
          a = K(2.0) * dr * ds;
          b = urq - usq;
          b = ncr * ncs * rbr * SQRT(b * b + K(.04) * urq * usq);
          clrs = b * cclrs * (K(5.0) * diaqr2 + a + K(.5) * diaqs2) * diaqr3;
          clsr = b * cclsr * (K(5.0) * diaqs2 + a + K(.5) * diaqr2) * diaqs3;
        } else { clrs = K(0.0); clsr = K(0.0); }
        if (qg > thresq) {
          a = K(5.0) * diaqr2 + K(2.0) * dr * dg + K(.5) * diaqg2;
          b = urq - ugq;
          clrg = a * cclrg * diaqr3 * ncr * ncg * rbr * SQRT(b * b + K(.04) * urq * ugq);
        } else clrg = K(0.0);
        sink = clri + clrs + clrg;
        if (qr < sink) { a = qr / sink; clri = clri * a; clrs = clrs * a; clrg = clrg * a; }
      } else { clri = clrs = clrg = clir = clsr = K(0.0); }
    }
    if (qi > thresq) {
      if (t < C_T0) {
        clis = qs > thresq ? cclis * qs2b * ncs * qi : K(0.0);
        clig = qg > thresq ? cclig * qg2b * ncg * qi : K(0.0);
        sink = clir + clis + clig;
        if (qi < sink) { a = qi / sink; clir = clir * a; clis = clis * a; clig = clig * a; }
      } else { clis = K(0.0); clig = K(0.0); }
    } else { clis = K(0.0); clig = K(0.0); }
    if (qs > thresq) {
      if (qg > thresq) {
        a = K(5.0) * diaqs2 + K(2.0) * ds * dg + K(.5) * diaqg2;
        b = usq - ugq;
        if (t < C_T0) clsg = a * cclsg * diaqs3 * ncs * ncg * rbr * SQRT(b * b + K(.04) * usq * ugq);
        else clsg = a * cclsg3 * diaqs3 * ncs * ncg * rbr * SQRT(b * b + K(.04) * usq * ugq);
      } else clsg = K(0.0);
      sink = clsr + clsg;
      if (qs < sink) { a = qs / sink; clsr = clsr * a; clsg = clsg * a; }
    } else clsg = K(0.0);
    W(CLCR)[p] = clcr; W(CLCS)[p] = clcs; W(CLCG)[p] = clcg; W(CLRI)[p] = clri; W(CLRS)[p] = clrs;
    W(CLRG)[p] = clrg; W(CLIR)[p] = clir; W(CLIS)[p] = clis; W(CLIG)[p] = clig; W(CLSR)[p] = clsr;
    W(CLSG)[p] = clsg; W(ECS)[p] = ecs;
  }
}

/* s_prodctwg, prodctwg.f90:206-209 constants, 401-471 (cphopt=2): wet growth
 * of graupel */
static void or_prodctwg(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R cc2dtn = K(-2.0) * C_CC * dtb, eigiv = K(1.0) / C_EIG, esgiv = K(1.0) / C_ESG;
  MLOOP {
    size_t p = PIX;
    if (s->qgp[p] > thresq && W(TCEL)[p] < C_T0CEL) {
      R tcel = W(TCEL)[p], rbr = s->rbr[p];
      R pgdry = W(CLCG)[p] + W(CLRG)[p] + W(CLIG)[p] + W(CLSG)[p];
      R lfice = W(LF)[p] + C_CW * tcel;
      R cligw = eigiv * W(CLIG)[p], clsgw = esgiv * W(CLSG)[p];
      R pg = cc2dtn * W(VNTG)[p] * (W(LV)[p] * W(DV)[p] * rbr * W(QVSST0)[p] + W(KP)[p] * tcel) /
                 (lfice * rbr) +
             (cligw + clsgw) * (K(1.0) - C_CI * tcel / lfice);
      W(PGWET)[p] = pg;
      if (pg > K(0.0) && pg < pgdry) {
        W(CLIG)[p] = cligw;
        W(CLSG)[p] = clsgw;
        R sink = W(CLIR)[p] + W(CLIS)[p] + W(CLIG)[p];
        if (s->qip[p] < sink) {
          R a = s->qip[p] / sink;
          W(CLIR)[p] = W(CLIR)[p] * a; W(CLIS)[p] = W(CLIS)[p] * a; W(CLIG)[p] = W(CLIG)[p] * a;
        }
        sink = W(CLSR)[p] + W(CLSG)[p];
        if (s->qsp[p] < sink) {
          R a = s->qsp[p] / sink;
          W(CLSR)[p] = W(CLSR)[p] * a; W(CLSG)[p] = W(CLSG)[p] * a;
        }
      } else {
        W(PGWET)[p] = K(-1.0);
      }
    } else {
      W(PGWET)[p] = K(-1.0);
    }
  }
}

/* s_distrpg, distrpg.f90:148-149, 274-314 (cphopt=2): rain-snow collection
 * split between snow and graupel */
static void or_distrpg(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R rhos2 = C_RHOS * C_RHOS, rhow2 = C_RHOW * C_RHOW;
  MLOOP {
    size_t p = PIX;
    if (s->qsp[p] > thresq) {
      if (W(T)[p] < C_T0) {
        R a = W(DIAQS)[p] * W(DIAQS)[p];
        R b = W(DIAQR)[p] * W(DIAQR)[p];
        a = rhos2 * a * a * a;
        b = rhow2 * b * b * b;
        R alpha = a / (a + b), alpha1 = K(1.0) - alpha;
        W(CLRSG)[p] = alpha1 * W(CLRS)[p];
        W(CLRS)[p] = alpha * W(CLRS)[p];
        W(CLSR)[p] = alpha1 * W(CLSR)[p];
      } else {
        W(CLRSG)[p] = W(CLRS)[p];
        W(CLRS)[p] = K(0.0);
      }
    } else {
      W(CLRSG)[p] = K(0.0);
    }
  }
}

/* s_aggregat, aggregat.f90:188, 399-419 (cphopt=2): cloud-ice number loss */
static void or_aggregat(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R cagin = C_R0 * EXP(K(3.0) * LOG(K(.5) * C_AUI * C_EII * C_XL / C_RHOI * dtb));
  MLOOP {
    size_t p = PIX;
    if (s->qip[p] > thresq)
      W(AGIN)[p] = s->rbr[p] * s->ncip[p] * s->qip[p] * EXP(ONED3 * LOG(cagin * s->rbv[p]));
    else
      W(AGIN)[p] = K(0.0);
  }
}

/* s_melting, melting.f90:174, 280-360 */
static void or_melting(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R cc2dt = K(2.0) * C_CC * dtb;
  MLOOP {
    size_t p = PIX;
    R tcel = W(TCEL)[p];
    R cmlxr1 = cc2dt * s->rbv[p] * (W(KP)[p] * tcel + W(LV)[p] * W(DV)[p] * s->rbr[p] * W(QVSST0)[p]);
    R cmlxr2 = C_CW * tcel;
    W(MLIC)[p] = (s->qip[p] > thresq && tcel >= C_T0CEL) ? s->qip[p] : K(0.0);
    if (s->qsp[p] > thresq && tcel >= C_T0CEL)
      W(MLSR)[p] = rmax((cmlxr1 * W(VNTS)[p] + cmlxr2 * (W(CLCS)[p] + W(CLRS)[p])) / W(LF)[p], K(0.0));
    else
      W(MLSR)[p] = K(0.0);
    if (s->qgp[p] > thresq && tcel >= C_T0CEL)
      W(MLGR)[p] = rmax((cmlxr1 * W(VNTG)[p] + cmlxr2 * (W(CLCG)[p] + W(CLRG)[p])) / W(LF)[p], K(0.0));
    else
      W(MLGR)[p] = K(0.0);
  }
}

/* s_depsit, depsit.f90:231-233, 429-600: deposition/evaporation */
static void or_depsit(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R ccdtb4 = K(2.0) * C_CC * dtb;
  const R t27311 = C_T0 - K(.01);
  koenig_table();
  MLOOP {
    size_t p = PIX;
    R t = W(T)[p];
    if (t > C_TLOW) {
      R qv = s->qvp[p], qvsw = W(QVSW)[p], qvsi = W(QVSI)[p];
      R qvssw = qv - qvsw, qvssi = qv - qvsi;
      R a = K(1.0) / (C_RV * W(KP)[p] * t * t);
      R b = W(DV)[p] * s->rbr[p];
      R cc = ccdtb4 * s->rbv[p];
      R ls = W(LS)[p], lv = W(LV)[p];
      R gi = K(1.0) / (ls * ls * a + K(1.0) / (b * qvsi));
      R cvdvx1 = cc * (qv / qvsw - K(1.0)) / (lv * lv * a + K(1.0) / (b * qvsw));
      R cvdvx2 = ccdtb4 * W(DV)[p] * W(QVSST0)[p];
      R cvdvx3 = cc * (qv / qvsi - K(1.0));
      R cvdvx4 = a * ls * W(LF)[p];
      R vdvr, vdvi, vdvs, vdvg;
      if (s->qrp[p] > thresq && qv < qvsw) vdvr = rmax(cvdvx1 * W(VNTR)[p], qvssw);
      else vdvr = K(0.0);
      if (s->qip[p] > thresq && t < t27311) {
        int itsc = (int)(-W(TCEL)[p]);
        vdvi = g_ckoe[itsc] * (qv - qvsi) / (qvsw - qvsi) * s->ncip[p] *
               EXP(g_pkoe[itsc] * LOG(W(MI)[p])) * dtb;
      } else vdvi = K(0.0);
      if (s->qsp[p] > thresq) {
        if (t > C_T0) vdvs = W(MLSR)[p] > K(0.0) ? cvdvx2 * W(VNTS)[p] : cvdvx1 * W(VNTS)[p];
        else vdvs = gi * (cvdvx3 * W(VNTS)[p] - cvdvx4 * W(CLCS)[p]);
      } else vdvs = K(0.0);
      if (s->qgp[p] > thresq) {
        if (t > C_T0) vdvg = W(MLGR)[p] > K(0.0) ? cvdvx2 * W(VNTG)[p] : cvdvx1 * W(VNTG)[p];
        else vdvg = gi * (cvdvx3 * W(VNTG)[p] - cvdvx4 * W(CLCG)[p]);
      } else vdvg = K(0.0);
      R sink = vdvi + vdvs + vdvg;
      if ((qvssi < sink && qvssi > K(0.0)) || (qvssi > sink && qvssi < K(0.0))) {
        R h = qvssi / sink;
        vdvi = vdvi * h; vdvs = vdvs * h; vdvg = vdvg * h;
      }
      W(VDVR)[p] = vdvr; W(VDVI)[p] = vdvi; W(VDVS)[p] = vdvs; W(VDVG)[p] = vdvg;
    } else {
      W(VDVR)[p] = K(0.0); W(VDVI)[p] = K(0.0); W(VDVS)[p] = K(0.0); W(VDVG)[p] = K(0.0);
    }
  }
}

/* s_convers, convers.f90:201-216, 481-588 (cphopt=2): autoconversion */
static void or_convers(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R ms05 = K(.5) * C_MS0;
  const R qccm = ONED6 * C_RHOW * C_CC * C_DIAQCM * C_DIAQCM * C_DIAQCM;
  const R diaqs0 = EXP(ONED3 * LOG(K(6.0) * C_MS0 / (C_CC * C_RHOS)));
  const R cagin = C_R0 * EXP(K(3.0) * LOG(K(.5) * ONED3 * C_AUI * C_EII * C_XL / C_RHOI * dtb));
  const R ccncr = K(.104) * C_G * C_ECC * EXP(-ONED3 * LOG(C_RHOW)) * dtb;
  R ccnsgn = K(1.0) / (C_RHOG - C_RHOS);
  const R ccnsg = C_RHOG * ccnsgn;
  MLOOP {
    size_t p = PIX;
    R qc = s->qcp[p], t = W(T)[p];
    if (qc > thresq) {
      R qcm = qccm * s->nccp[p];
      if (t > C_TLOW && qc >= qcm)
        W(CNCR)[p] = ccncr * s->rbr[p] * EXP(-ONED3 * LOG(s->nccp[p])) * EXP(SEVND3 * LOG(qc)) / W(MU)[p];
      else
        W(CNCR)[p] = K(0.0);
    } else {
      W(CNCR)[p] = K(0.0);
    }
    if (t < C_T0) {
      R qi = s->qip[p];
      if (qi > thresq) {
        R cn = s->rbr[p] * EXP(ONED3 * LOG(cagin * s->rbv[p])) * qi * qi / LOG(diaqs0 / W(DIAQI)[p]);
        if (W(VDVI)[p] >= K(0.0)) {
          R mi = W(MI)[p];
          if (mi < ms05) cn = cn + mi / (C_MS0 - mi) * W(VDVI)[p];
          else cn = cn + (W(VDVI)[p] + (K(1.0) - ms05 / mi) * qi);
        }
        W(CNIS)[p] = rmin(cn, qi);
      } else {
        W(CNIS)[p] = K(0.0);
      }
      if (s->qsp[p] > thresq && W(VDVS)[p] > K(0.0) && W(VDVS)[p] < W(CLCS)[p])
        W(CNSG)[p] = rmin(ccnsg * W(CLCS)[p], s->qsp[p]);
      else
        W(CNSG)[p] = K(0.0);
    } else {
      W(CNIS)[p] = K(0.0);
      W(CNSG)[p] = K(0.0);
    }
  }
}

/* s_nuc2nd, nuc2nd.f90:127-128, 210-273: secondary ice (Hallett-Mossop) */
static void or_nuc2nd(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R mi0352 = K(.5) * K(3.5e8) * C_MI0, mi0353 = ONED3 * K(3.5e8) * C_MI0;
  MLOOP {
    size_t p = PIX;
    R t = W(T)[p], rbv = s->rbv[p];
    if (W(PGWET)[p] > K(0.0)) {
      if (t > K(270.16)) W(SPSI)[p] = K(0.0);
      else if (t > K(268.16) && t <= K(270.16)) W(SPSI)[p] = (K(270.16) - t) * mi0352 * rbv * W(CLCS)[p];
      else if (t > K(265.16) && t <= K(268.16)) W(SPSI)[p] = (t - K(265.16)) * mi0353 * rbv * W(CLCS)[p];
      else W(SPSI)[p] = K(0.0);
      W(SPGI)[p] = K(0.0);
    } else {
      if (t > K(270.16)) { W(SPSI)[p] = K(0.0); W(SPGI)[p] = K(0.0); }
      else if (t > K(268.16) && t <= K(270.16)) {
        R a = (K(270.16) - t) * mi0352 * rbv;
        W(SPSI)[p] = a * W(CLCS)[p]; W(SPGI)[p] = a * W(CLCG)[p];
      } else if (t > K(265.16) && t <= K(268.16)) {
        R a = (t - K(265.16)) * mi0353 * rbv;
        W(SPSI)[p] = a * W(CLCS)[p]; W(SPGI)[p] = a * W(CLCG)[p];
      } else { W(SPSI)[p] = K(0.0); W(SPGI)[p] = K(0.0); }
    }
  }
}

/* s_freezing, freezing.f90:145-148, 266-302 (cphopt=2): rain -> graupel */
static void or_freezing(const or_cfg *c, or_src *s, R dtb)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R tclow = C_TLOW - C_T0;
  const R cfrrg = K(20.0) * K(100.0) * C_CC * C_CC * C_RHOW * dtb;
  MLOOP {
    size_t p = PIX;
    R qr = s->qrp[p], tcel = W(TCEL)[p];
    if (qr > thresq && tcel < C_T0CEL) {
      if (tcel <= tclow) W(FRRG)[p] = qr;
      else {
        R d = W(DIAQR)[p];
        R d3 = d * d * d;
        W(FRRG)[p] = rmin(qr, cfrrg * d3 * d3 * (EXP(K(-.66) * tcel) - K(1.0)) * s->ncrp[p]);
      }
    } else {
      W(FRRG)[p] = K(0.0);
    }
  }
}

/* s_shedding, shedding.f90:211-267 */
static void or_shedding(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  MLOOP {
    size_t p = PIX;
    R t = W(T)[p];
    W(SHSR)[p] = (s->qsp[p] > thresq && t >= C_T0) ? W(CLCS)[p] + W(CLRS)[p] : K(0.0);
    if (s->qgp[p] > thresq) {
      if (t >= C_T0) W(SHGR)[p] = W(CLCG)[p] + W(CLRG)[p];
      else if (W(PGWET)[p] > K(0.0))
        W(SHGR)[p] = (W(CLCG)[p] + W(CLRG)[p] + W(CLIG)[p] + W(CLSG)[p]) - W(PGWET)[p];
      else W(SHGR)[p] = K(0.0);
    } else {
      W(SHGR)[p] = K(0.0);
    }
  }
}

/* s_more0q, more0q.f90:490-646 (evapor_opt=1): scale sinks so that no
 * future hydrometeor goes negative */
static void or_more0q(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
#define S(n) W(n)[p]
  MLOOP {
    size_t p = PIX;
    R sink, h;
    if (s->qcp[p] > thresq) {
      sink = S(NUCI) + S(CLCR) + S(CLCS) + S(CLCG) + S(CNCR) - S(MLIC);
      if (s->qcf[p] < sink) {
        h = s->qcf[p] / sink;
        S(NUCI) = S(NUCI) * h; S(CLCR) = S(CLCR) * h; S(CLCS) = S(CLCS) * h; S(CLCG) = S(CLCG) * h;
        S(CNCR) = S(CNCR) * h; S(MLIC) = S(MLIC) * h;
      }
    }
    if (s->qip[p] > thresq) {
      sink = S(CLIR) + S(CLIS) + S(CLIG) + S(CNIS) + S(MLIC) - S(VDVI) - S(NUVI) - S(NUCI) - S(SPSI) - S(SPGI);
      if (s->qif[p] < sink) {
        h = s->qif[p] / sink;
        S(CLIR) = S(CLIR) * h; S(CLIS) = S(CLIS) * h; S(CLIG) = S(CLIG) * h; S(CNIS) = S(CNIS) * h;
        S(MLIC) = S(MLIC) * h; S(VDVI) = S(VDVI) * h; S(NUVI) = S(NUVI) * h; S(NUCI) = S(NUCI) * h;
        S(SPSI) = S(SPSI) * h; S(SPGI) = S(SPGI) * h;
      }
    }
    if (s->qrp[p] > thresq) {
      sink = S(CLRI) + S(CLRS) + S(CLRG) + S(CLRSG) + S(FRRG) - S(VDVR) - S(CLCR) - S(CNCR) - S(MLSR) -
             S(MLGR) - S(SHSR) - S(SHGR);
      if (s->qrf[p] < sink) {
        h = s->qrf[p] / sink;
        S(CLRI) = S(CLRI) * h; S(CLRS) = S(CLRS) * h; S(CLRG) = S(CLRG) * h; S(CLRSG) = S(CLRSG) * h;
        S(FRRG) = S(FRRG) * h; S(VDVR) = S(VDVR) * h; S(CLCR) = S(CLCR) * h; S(CNCR) = S(CNCR) * h;
        S(MLSR) = S(MLSR) * h; S(MLGR) = S(MLGR) * h; S(SHSR) = S(SHSR) * h; S(SHGR) = S(SHGR) * h;
      }
    }
    if (s->qsp[p] > thresq) {
      sink = S(CLSR) + S(CLSG) + S(CNSG) + S(SPSI) + S(MLSR) + S(SHSR) - S(VDVS) - S(CLCS) - S(CLRS) -
             S(CLIS) - S(CNIS);
      if (s->qsf[p] < sink) {
        h = s->qsf[p] / sink;
        S(CLSR) = S(CLSR) * h; S(CLSG) = S(CLSG) * h; S(CNSG) = S(CNSG) * h; S(SPSI) = S(SPSI) * h;
        S(MLSR) = S(MLSR) * h; S(SHSR) = S(SHSR) * h; S(VDVS) = S(VDVS) * h; S(CLCS) = S(CLCS) * h;
        S(CLRS) = S(CLRS) * h; S(CLIS) = S(CLIS) * h; S(CNIS) = S(CNIS) * h;
      }
    }
    if (s->qgp[p] > thresq) {
      sink = S(SPGI) + S(MLGR) + S(SHGR) - S(VDVG) - S(CLRI) - S(CLIR) - S(CLSR) - S(CLCG) - S(CLRG) -
             S(CLIG) - S(CLSG) - S(CLRSG) - S(CNSG) - S(FRRG);
      if (s->qgf[p] < sink) {
        h = s->qgf[p] / sink;
        S(SPGI) = S(SPGI) * h; S(MLGR) = S(MLGR) * h; S(SHGR) = S(SHGR) * h; S(VDVG) = S(VDVG) * h;
        S(CLRI) = S(CLRI) * h; S(CLIR) = S(CLIR) * h; S(CLSR) = S(CLSR) * h; S(CLCG) = S(CLCG) * h;
        S(CLRG) = S(CLRG) * h; S(CLIG) = S(CLIG) * h; S(CLSG) = S(CLSG) * h; S(CLRSG) = S(CLRSG) * h;
        S(CNSG) = S(CNSG) * h; S(FRRG) = S(FRRG) * h;
      }
    }
  }
}

/* s_newblk, newblk.f90:362-364, 985-1129 (cphopt=2): apply the sources to
 * the future level, with latent heating */
static void or_newblk(const or_cfg *c, or_src *s)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R mi0iv = K(1.0) / C_MI0, ms0iv = K(1.0) / C_MS0;
  MLOOP {
    size_t p = PIX;
    R cppiv = K(1.0) / (C_CP * s->pi[p]);
    R clix = S(CLIR) + S(CLIS) + S(CLIG);
    R spxi = S(SPSI) + S(SPGI);
    R cxcr = S(CLCR) + S(CNCR);
    R cxsg = S(CLSR) + S(CLSG) + S(CNSG);
    R nmci = S(NUCI) - S(MLIC);
    R mssr = S(MLSR) + S(SHSR);
    R msgr = S(MLGR) + S(SHGR);
    R nuvdvx = S(VDVI) + S(VDVS) + S(VDVG) + S(NUVI);
    R clnmci = S(CLCS) + S(CLCG) + nmci;
    R cfmsxr = S(CLRI) + S(CLRS) + S(CLRG) + S(CLRSG) + S(FRRG) - mssr - msgr;
    s->ptpf[p] = s->ptpf[p] + (S(VDVR) * S(LV) + (clnmci + cfmsxr) * S(LF) + nuvdvx * S(LS)) * cppiv;
    R qvsink = nuvdvx + S(VDVR);
    R qxsink = cfmsxr - cxcr - S(VDVR);
    if (s->qrf[p] < qxsink) {
      qxsink = cxcr + clnmci + (qxsink - s->qrf[p]);
      s->qrf[p] = K(0.0);
    } else {
      s->qrf[p] = s->qrf[p] - qxsink;
      qxsink = cxcr + clnmci;
    }
    if (s->qcf[p] < qxsink) {
      R dqv = qxsink - s->qcf[p];
      s->ptpf[p] = s->ptpf[p] + cppiv * dqv * S(LV);
      qvsink = qvsink + dqv;
      s->qcf[p] = K(0.0);
    } else {
      s->qcf[p] = s->qcf[p] - qxsink;
    }
    qxsink = clix - spxi - nmci + S(CNIS) - S(VDVI) - S(NUVI);
    if (s->qif[p] < qxsink) {
      R dqv = qxsink - s->qif[p];
      s->ptpf[p] = s->ptpf[p] + cppiv * dqv * S(LS);
      qvsink = qvsink + dqv;
      s->qif[p] = K(0.0);
    } else {
      s->qif[p] = s->qif[p] - qxsink;
    }
    s->qvf[p] = rmax(s->qvf[p] - qvsink, K(0.0));
    s->qsf[p] = rmax(s->qsf[p] - (cxsg + mssr + S(SPSI) - S(VDVS) - S(CLCS) - S(CLRS) - S(CLIS) - S(CNIS)),
                     K(0.0));
    s->qgf[p] = rmax(s->qgf[p] + (cxsg - msgr - S(SPGI) + S(VDVG) + S(CLRI) + S(CLIR) + S(CLCG) + S(CLRG) +
                                  S(CLIG) + S(CLRSG) + S(FRRG)),
                     K(0.0));
    if (s->qcp[p] > thresq) s->ncif[p] = s->ncif[p] + S(NUCI) * s->nccp[p] / s->qcp[p];
    if (s->qip[p] > thresq) {
      if (S(VDVI) < K(0.0))
        s->ncif[p] = s->ncif[p] - (S(AGIN) + (clix + S(MLIC) - S(VDVI)) / S(MI) - (spxi + S(NUVI)) * mi0iv +
                                   S(CNIS) * ms0iv);
      else
        s->ncif[p] = s->ncif[p] - (S(AGIN) + (clix + S(MLIC)) / S(MI) - (spxi + S(NUVI)) * mi0iv +
                                   S(CNIS) * ms0iv);
    } else {
      s->ncif[p] = s->ncif[p] + (spxi + S(NUVI)) * mi0iv;
    }
  }
#undef S
}

/* s_srcblk, srcblk.f90:488-617: the source-term chain, in the reference's
 * call order (nuc2nd_opt=1, evapor_opt=1, cphopt=2). */
void or_srcblk(const or_cfg *c, or_src *s, R dtb)
{
  or_setblk(c, s);
  or_nuc1stv(c, s);
  or_nuc1stc(c, s, dtb);
  or_collect(c, s, dtb);
  or_prodctwg(c, s, dtb);
  or_distrpg(c, s);
  or_aggregat(c, s, dtb);
  or_melting(c, s, dtb);
  or_depsit(c, s, dtb);
  or_convers(c, s, dtb);
  or_nuc2nd(c, s);
  or_freezing(c, s, dtb);
  or_shedding(c, s);
  or_more0q(c, s);
  or_newblk(c, s);
}

/* s_upwqp, upwqp.f90:149-225 (advopt<=3): one upwind sedimentation
 * sub-step of a mixing ratio, with the surface precipitation rate and its
 * accumulation (the leapfrog accumulates dtp/2 per sub-step). */
void or_upwqp(const or_cfg *c, R dtp, const R *rbr, const R *rst, const R *uq, R *qpf, R *precip, R *qpflx)
{
  DECL_DIMS(c);
  const R dziv = (R)1 / (R)c->dz;
  const R dtp05 = K(.5) * dtp, dzvdt = dziv * dtp, rwiv05 = K(.5) / C_RHOW;
  const size_t n2 = NX * NY;
  MLOOP { size_t p = PIX; qpflx[p] = rbr[p] * uq[p] * qpf[p]; }
  OMP_FOR
  for (int k = 1; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        size_t p = PIX;
        qpf[p] = rmax(qpf[p] + (qpflx[p + n2] - qpflx[p]) / rst[p] * dzvdt, K(0.0));
      }
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni - 1; i++) {
      size_t q = (size_t)i + NX * (size_t)j;
      size_t top = q + n2 * (size_t)(nk - 2), below = q + n2 * (size_t)(nk - 3);
      qpf[top] = qpf[below];
      precip[q] = (qpflx[q] + qpflx[q + n2]) * rwiv05;
      precip[q + n2] = precip[q + n2] + precip[q] * dtp05;
    }
}

/* s_upwnp, upwnp.f90:132-176: the same for a number concentration */
void or_upwnp(const or_cfg *c, R dtp, const R *rbr, const R *rst, const R *un, R *ncf, R *ncflx)
{
  DECL_DIMS(c);
  const R dziv = (R)1 / (R)c->dz;
  const R dzvdt = dziv * dtp;
  const size_t n2 = NX * NY;
  MLOOP { size_t p = PIX; ncflx[p] = rbr[p] * un[p] * ncf[p]; }
  OMP_FOR
  for (int k = 1; k <= nk - 2; k++)
    for (int j = 1; j <= nj - 1; j++)
      for (int i = 1; i <= ni - 1; i++) {
        size_t p = PIX;
        ncf[p] = rmax(ncf[p] + (ncflx[p + n2] - ncflx[p]) / rst[p] * dzvdt, K(0.0));
      }
  for (int j = 1; j <= nj - 1; j++)
    for (int i = 1; i <= ni - 1; i++) {
      size_t q = (size_t)i + NX * (size_t)j;
      ncf[q + n2 * (size_t)(nk - 2)] = ncf[q + n2 * (size_t)(nk - 3)];
    }
}

/* s_fallblk, fallblk.f90:298-549 (flqcqi_opt=1, haiopt=0): CFL-limited
 * sub-step count from the domain minimum of dz*jcb/(U+eps) (chkfall's
 * MIN over ranks is the identity on one domain), then npstp sub-steps of
 * upwind sedimentation for qc, qi, nci, qr, qs, qg.  Returns npstp. */
int or_fallblk(const or_cfg *c, R dtb, const R *jcb, const R *rbr, const R *rst, const R *ucq,
               const R *urq, const R *uiq, const R *usq, const R *ugq, const R *uin, R *qcf, R *qrf,
               R *qif, R *qsf, R *qgf, R *ncif, R *prc, R *prr, R *pri, R *prs, R *prg, R *tmp1,
               R *dtp_out)
{
  DECL_DIMS(c);
  const R dz = (R)c->dz;
  R dtpc = dtb, dtpr = dtb, dtpi = dtb, dtps = dtb, dtpg = dtb;
  MLOOP_SERIAL {
    size_t p = PIX;
    R dzjcb = dz * jcb[p];
    dtpc = rmin(dzjcb / (ucq[p] + C_EPS), dtpc);
    dtpr = rmin(dzjcb / (urq[p] + C_EPS), dtpr);
    dtpi = rmin(dzjcb / (uiq[p] + C_EPS), dtpi);
    dtps = rmin(dzjcb / (usq[p] + C_EPS), dtps);
    dtpg = rmin(dzjcb / (ugq[p] + C_EPS), dtpg);
  }
  R dtp = rmin(rmin(rmin(rmin(dtpc, dtpr), dtpi), dtps), dtpg);
  int npstp;
  if (dtp < dtb) {
    npstp = (int)((dtb + K(.001)) / dtp) + 1;
    dtp = dtb / (R)npstp;
  } else {
    npstp = 1;
    dtp = dtb;
  }
  for (int s = 0; s < npstp; s++) {
    or_upwqp(c, dtp, rbr, rst, ucq, qcf, prc, tmp1);
    or_upwqp(c, dtp, rbr, rst, uiq, qif, pri, tmp1);
    or_upwnp(c, dtp, rbr, rst, uin, ncif, tmp1);
  }
  for (int s = 0; s < npstp; s++) {
    or_upwqp(c, dtp, rbr, rst, urq, qrf, prr, tmp1);
    or_upwqp(c, dtp, rbr, rst, usq, qsf, prs, tmp1);
    or_upwqp(c, dtp, rbr, rst, ugq, qgf, prg, tmp1);
  }
  if (dtp_out) *dtp_out = dtp;
  return npstp;
}

/* s_siadjst, siadjst.f90:149-269: ice saturation adjustment below tlow,
 * applied twice */
void or_siadjst(const or_cfg *c, const R *ptbr, const R *pi, const R *p, R *ptp, R *qv, R *qi, R *nci)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  const R cwmci = C_CW - C_CI, mi0iv = K(1.0) / C_MI0;
  MLOOP {
    size_t q = PIX;
    R t = (ptbr[q] + ptp[q]) * pi[q];
    if (t > C_TLOW) continue;
    for (int pass = 0; pass < 2; pass++) {
      if (pass == 1) {
        t = (ptbr[q] + ptp[q]) * pi[q];
        if (!(t <= C_TLOW)) break;
      }
      R tcel = t - C_T0;
      R a = K(1.0) / (t - K(7.66));
      R b = a * tcel;
      R esi = C_ES0 * EXP(K(21.875) * b);
      R qvsi = C_EPSVA * esi / (p[q] - esi);
      if (qi[q] > thresq || qv[q] > qvsi) {
        R lscpi = (C_LV0 * EXP((K(.167) + K(3.67e-4) * t) * LOG(C_T0 / t)) + (C_LF0 + cwmci * tcel)) /
                  (C_CP * pi[q]);
        R dqi = (qvsi - qv[q]) / (K(1.0) + K(21.875) * a * (K(1.0) - b) * qvsi * lscpi * pi[q]);
        if (qi[q] > dqi) {
          if (qi[q] > thresq) nci[q] = nci[q] - dqi * nci[q] / qi[q];
          else nci[q] = nci[q] - dqi * mi0iv;
          ptp[q] = ptp[q] - dqi * lscpi;
          qv[q] = qv[q] + dqi;
          qi[q] = qi[q] - dqi;
        } else {
          nci[q] = K(0.0);
          ptp[q] = ptp[q] - qi[q] * lscpi;
          qv[q] = qv[q] + qi[q];
          qi[q] = K(0.0);
        }
      }
    }
  }
}

/* s_swadjst, swadjst.f90:237-310 (cphopt<=3): water saturation adjustment,
 * applied twice when the first pass found cloud or supersaturation */
void or_swadjst(const or_cfg *c, const R *ptbr, const R *pi, const R *p, R *ptp, R *qv, R *qc)
{
  DECL_DIMS(c);
  const R thresq = (R)c->thresq;
  MLOOP {
    size_t q = PIX;
    for (int pass = 0; pass < 2; pass++) {
      R t = (ptbr[q] + ptp[q]) * pi[q];
      R a = K(1.0) / (t - K(35.86));
      R b = a * (t - C_T0);
      R esw = C_ES0 * EXP(K(17.269) * b);
      R qvsw = C_EPSVA * esw / (p[q] - esw);
      if (!(qc[q] > thresq || qv[q] > qvsw)) break;
      R lvcpi = C_LV0 * EXP((K(.167) + K(3.67e-4) * t) * LOG(C_T0 / t)) / (C_CP * pi[q]);
      R dqc = (qvsw - qv[q]) / (K(1.0) + K(17.269) * a * (K(1.0) - b) * qvsw * lvcpi * pi[q]);
      if (qc[q] > dqc) {
        ptp[q] = ptp[q] - dqc * lvcpi;
        qv[q] = qv[q] + dqc;
        qc[q] = qc[q] - dqc;
      } else {
        ptp[q] = ptp[q] - qc[q] * lvcpi;
        qv[q] = qv[q] + qc[q];
        qc[q] = K(0.0);
      }
    }
  }
}

/* s_satadjst, satadjst.f90:154-187 (cphopt=2): getexner with this level's
 * p', then siadjst and swadjst */
void or_satadjst_level(const or_cfg *c, const R *pbr, const R *ptbr, const R *pp, R *ptp, R *q, R *nci,
                       R *pi, R *p)
{
  const size_t n = (size_t)nfield(c);
  or_getexner(c, pbr, pp, pi, p);
  or_siadjst(c, ptbr, pi, p, ptp, q, q + 3 * n, nci);
  or_swadjst(c, ptbr, pi, p, ptp, q, q + n);
}

/* s_cloudphy (cloudphy.f90:277-338) -> s_coldblk (coldblk.f90:321-481) for
 * cphopt=2, haiopt=0, advopt<=3: past level in, future level updated. */
int or_cloudphy(const or_cfg *c, or_state *s, R dtb, int *npstp_out)
{
  const size_t n = (size_t)nfield(c);
  const int P = 0, F = 2;
  R *w[MW_NW];
  R *blk = (R *)calloc((size_t)MW_NW * n, sizeof(R));
  if (!blk) return -1;
  for (int m = 0; m < MW_NW; m++) w[m] = blk + (size_t)m * n;
  const R dtbs = K(2.0) * dtb;
  R *qp = s->q[P], *qf = s->q[F];
  or_getvdens(c, s->rbr, w[MW_RBV]);
  or_getexner(c, s->pbr, s->pp[P], w[MW_PI], w[MW_P]);
  or_diagnw(c, s->rbr, qp + n, qp + 2 * n, w[MW_NCC], w[MW_NCR]);
  or_diagnsg(c, s->rbr, w[MW_RBV], qp + 4 * n, qp + 5 * n, w[MW_NCS], w[MW_NCG]);
  or_adjstnci(c, qf + 3 * n, s->nci[F]);
  or_termblk(c, w[MW_RBV], qp + n, qp + 2 * n, qp + 3 * n, qp + 4 * n, qp + 5 * n, w[MW_NCR], w[MW_NCS],
             w[MW_NCG], w[MW_UCQ], w[MW_URQ], w[MW_UIQ], w[MW_USQ], w[MW_UGQ], w[MW_UCN], w[MW_URN], w[MW_UIN],
             w[MW_USN], w[MW_UGN]);
  or_src src;
  src.ptbr = s->ptbr; src.rbr = s->rbr; src.rbv = w[MW_RBV]; src.pi = w[MW_PI]; src.p = w[MW_P];
  src.ptpp = s->ptp[P]; src.qvp = qp; src.qcp = qp + n; src.qrp = qp + 2 * n; src.qip = qp + 3 * n;
  src.qsp = qp + 4 * n; src.qgp = qp + 5 * n;
  src.nccp = w[MW_NCC]; src.ncrp = w[MW_NCR]; src.ncip = s->nci[P]; src.ncsp = w[MW_NCS]; src.ncgp = w[MW_NCG];
  src.urq = w[MW_URQ]; src.usq = w[MW_USQ]; src.ugq = w[MW_UGQ];
  src.ptpf = s->ptp[F]; src.qvf = qf; src.qcf = qf + n; src.qrf = qf + 2 * n; src.qif = qf + 3 * n;
  src.qsf = qf + 4 * n; src.qgf = qf + 5 * n; src.ncif = s->nci[F];
  for (int m = 0; m < MW_NW; m++) src.w[m] = w[m];
  or_srcblk(c, &src, dtbs);
  or_adjstnci(c, qf + 3 * n, s->nci[F]);
  const size_t n2 = (size_t)(c->ni + 2) * (size_t)(c->nj + 2);
  int np = or_fallblk(c, dtbs, s->jcb, s->rbr, s->rst, w[MW_UCQ], w[MW_URQ], w[MW_UIQ], w[MW_USQ], w[MW_UGQ],
                      w[MW_UIN], qf + n, qf + 2 * n, qf + 3 * n, qf + 4 * n, qf + 5 * n, s->nci[F],
                      s->prwtr, s->prwtr + 2 * n2, s->price, s->price + 2 * n2, s->price + 4 * n2,
                      w[MW_FLX], NULL);
  or_adjstnci(c, qf + 3 * n, s->nci[F]);
  if (npstp_out) *npstp_out = np;
  free(blk);
  return 0;
}

/* s_cloudphy on given arrays (the tier-2 entry point's counterpart): qp, qf
 * hold qv, qc, qr, qi, qs, qg one field after another; the future level
 * holds steps' result on entry. */
int or_cloudphy_a(const or_cfg *c, R dtb, R *jcb, R *pbr, R *ptbr, R *rbr, R *rst, R *ppp, R *ptpp, R *qp,
                  R *ncip, R *ptpf, R *qf, R *ncif, R *prwtr, R *price)
{
  or_state st;
  memset(&st, 0, sizeof st);
  st.jcb = jcb;
  st.pbr = pbr;
  st.ptbr = ptbr;
  st.rbr = rbr;
  st.rst = rst;
  st.pp[0] = ppp;
  st.ptp[0] = ptpp;
  st.q[0] = qp;
  st.nci[0] = ncip;
  st.ptp[2] = ptpf;
  st.q[2] = qf;
  st.nci[2] = ncif;
  st.prwtr = prwtr;
  st.price = price;
  return or_cloudphy(c, &st, dtb, NULL);
}

static void micro_hook(const or_cfg *c, or_state *s, R dtb, R *qallp)
{
  (void)qallp;
  or_cloudphy(c, s, dtb, NULL);
}

static void satadj_hook(const or_cfg *c, or_state *s, R dtb, R *unused)
{
  (void)dtb;
  (void)unused;
  const size_t n = (size_t)nfield(c);
  R *pi = (R *)calloc(2 * n, sizeof(R));
  if (!pi) return;
  or_satadjst_level(c, s->pbr, s->ptbr, s->pp[1], s->ptp[1], s->q[1], s->nci[1], pi, pi + n);
  or_satadjst_level(c, s->pbr, s->ptbr, s->pp[2], s->ptp[2], s->q[2], s->nci[2], pi, pi + n);
  free(pi);
}

/* Array-block entry for checking or_srcblk against the reference's s_srcblk
 * with the same layout as oracle/ref_harness.f90:ref_srcblk:
 *   inp: ptbr,rbr,rbv,pi,p,ptpp,qvp,qcp,qrp,qip,qsp,qgp,nccp,ncrp,ncip,ncsp,
 *        ncgp,urq,usq,ugq (20 fields),
 *   fut: ptpf,qvf,qcf,qrf,qif,qsf,qgf,ncif,
 *   wrk: the reference's work arrays in its argument order (61 fields) plus
 *        one scratch field. */
void or_srcblk_block(const or_cfg *c, R dtb, const R *inp, R *fut, R *wrk)
{
  const size_t n = (size_t)nfield(c);
  or_src s;
  const R **ip[20] = {&s.ptbr, &s.rbr, &s.rbv, &s.pi, &s.p, &s.ptpp, &s.qvp, &s.qcp, &s.qrp, &s.qip,
                      &s.qsp, &s.qgp, &s.nccp, &s.ncrp, &s.ncip, &s.ncsp, &s.ncgp, &s.urq, &s.usq, &s.ugq};
  for (int f = 0; f < 20; f++) *ip[f] = inp + (size_t)f * n;
  R **fp[8] = {&s.ptpf, &s.qvf, &s.qcf, &s.qrf, &s.qif, &s.qsf, &s.qgf, &s.ncif};
  for (int f = 0; f < 8; f++) *fp[f] = fut + (size_t)f * n;
  static const int order[61] = {
      MW_T, MW_TCEL, MW_QVSST0, MW_QVSW, MW_QVSI, MW_LV, MW_LS, MW_LF, MW_KP, MW_MU, MW_DV, MW_MI,
      MW_DIAQC, MW_DIAQR, MW_DIAQI, MW_DIAQS, MW_DIAQG, MW_VNTR, MW_VNTS, MW_VNTG,
      MW_NUVI, MW_NUCI, MW_CLCR, MW_CLCS, MW_CLCG, MW_CLRI, MW_CLRS, MW_CLRG,
      MW_CLIR, MW_CLIS, MW_CLIG, MW_CLSR, MW_CLSG, MW_CLRSG, -1,
      -1, -1, -1, -1, -1, MW_AGIN, -1,
      MW_VDVR, MW_VDVI, MW_VDVS, MW_VDVG, MW_CNCR, MW_CNIS, MW_CNSG, -1,
      MW_SPSI, MW_SPGI, MW_MLIC, MW_MLSR, MW_MLGR, MW_FRRG, -1,
      MW_SHSR, MW_SHGR, MW_PGWET, MW_ECS};
  for (int m = 0; m < MW_NW; m++) s.w[m] = wrk + 61 * n;   /* arrays the reference keeps elsewhere */
  for (int f = 0; f < 61; f++)
    if (order[f] >= 0) s.w[order[f]] = wrk + (size_t)f * n;
  s.w[MW_FLX] = wrk + 61 * n;
  or_srcblk(c, &s, dtb);
}
