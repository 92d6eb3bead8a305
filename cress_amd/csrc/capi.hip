// C ABI (include/cress_hip.h) and the host-side orchestration of one CReSS
// large step on one MI355X: s_intgdrv (intgdrv.f90:832-1133) with
// fdmdrv/forcedrv/smalldrv/hevi/steps/cloudphy/timeflt/satadjst/shift2nd
// and the three-level rotation of swp2nxt done by renaming buffers.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "common.h"
#include "diag.h"
#include "dyn.h"
#include "setup.h"
#include "micro.h"
#include "obc.h"
#include "turb.h"

#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <utility>

using namespace cress;

#define HIPCHK(x)                                  \
  do {                                             \
    hipError_t e_ = (x);                           \
    if (e_ != hipSuccess) {                        \
      fprintf(stderr, "cress: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return CRESS_EHIP;                           \
    }                                              \
  } while (0)

enum { NSTAGE = 8 };


// Kernel families timed individually when timing is on (HIP events on the
// context stream around every launch).
enum {
  KT_PREP, KT_FORCE_UV, KT_FORCE_W, KT_FORCE_P, KT_FORCE_S, KT_PHY2CNT, KT_AC_PTUV,
  KT_AC_COL, KT_AC_RING, KT_FILL, KT_VBC, KT_STEPS, KT_TIMEFLT, KT_COPY, KT_MICRO_SEED, KT_MICRO_SRC,
  KT_MICRO_FALL, KT_SATADJ, KT_OBC, KT_MLP_SELECT, KT_MLP_INFER, KT_N
};
constexpr int kDiagBlocks = 1024;   // blocks of the FP64 diagnostics' first pass

// Default launch shapes: the fastest measured at 1024x1024x80 (dyn.h).
static LaunchTune tune_default() {
  LaunchTune t;
  t.ptuv_by = 16;
  t.ptuv_slots = 0;
  t.ptuv_order = 1;
  t.col_by = 2;
  t.col_lds = 4096;
  t.fs_by = 4;
  t.fs_order = 2;
  t.ff_by = 4;
  t.ff_order = 2;
  return t;
}

// The fused ptil/stepuv kernel for a grid (per-level tables or not) and a
// halo layout.
static auto ptuv_kernel(const Grid &g, int slots) {
  const unsigned upu = (1u << CRESS_RST8U) | (1u << CRESS_RST8V) | (1u << CRESS_RST8W) | (1u << CRESS_JCB);
  const bool u = (g.umask & upu) == upu;
  return slots ? (u ? k_ac_ptuv<true, 1> : k_ac_ptuv<false, 1>) : (u ? k_ac_ptuv<true, 0> : k_ac_ptuv<false, 0>);
}

static const char *kKernelNames[KT_N] = {
    "k_prep", "k_force_uv", "k_force_w", "k_force_p", "k_force_s", "k_phy2cnt", "k_ac_ptuv",
    "k_ac_col", "k_ac_ring", "k_bcycle", "k_vbc", "k_steps", "k_timeflt", "memcpy_d2d",
    "k_micro_seed", "k_micro_src", "k_micro_fall", "k_satadjst", "k_obc", "k_mlp_select", "k_mlp_infer"};
enum { MAXEV = 1024 };

struct cress_ctx {
  cress_config cfg;
  Grid g;
  Consts c;
  Opts o;
  long n;  // floats per 3-D field
  hipStream_t stream;
  // base state / metrics
  float *base[CRESS_NBASE];
  float *fc, *rbct;
  float *tab;  // per-level copies of the horizontally uniform base fields (g.tab)
  std::vector<float> tabh;  // host copy of tab
  // Thomas coefficients of the column kernel (AcCol::coef) for the dts of
  // slot s (accoef_dts[s]): slots 0 and 1 hold the cold-start and the full
  // substep, slot 2 any other dts a caller asks for; accoef_ok[s] = 0 when
  // the metrics are not uniform or the slot is stale
  float *accoef;
  float accoef_dts[3];
  int accoef_ok[3];
  double *diag;  // cress_diag: kDiagBlocks x DIAG_N partials, DIAG_N results, nk+1 heights
  LaunchTune tune;  // launch shapes (tune_default, cress_autotune)
  double *setupcol;   // cress_set_base's binary64 columns (SC_N x (nk+1)), or null
  int tune_ver;     // bumped whenever tune changes (captured ensemble graphs bake it in)
  // prognostics by physical slot; logical level l lives in slot (rot+l)%3
  float *u[3], *v[3], *w[3], *pp[3], *ptp[3], *q[3], *nci[3];
  int rot;
  float *prwtr, *price;  // precipitation (rate, accumulation)
  // scratch
  float *ufrc, *vfrc, *wfrc, *pfrc, *ptfrc, *qvfrc, *wc, *ptil, *rstxu, *rstxv, *rstxwc;
  float *ualt, *valt;  // second buffers of the future uf, vf during the acoustic loop
  float *sfrc;         // forcings of qc, qr, qi, qs, qg, nci (cphopt=2)
  MicroScratch ms;
  // open sides: dtb, dts of the last large step (the phase speeds phasev
  // leaves for the next one); obc_have=0 before any step
  float obc_dtb, obc_dts;
  int obc_have;
  std::vector<void *> allocs;
  // learned microphysics: 0 srcblk, 1 srcblk recording pairs into mlp,
  // 2 the network (cress_ctx_set_mlp); ml holds the per-step buffers
  struct cress_mlp *mlp;
  int mlp_mode;
  MlpStep ml;
  int timing;
  int tkern;  // timing of one kernel family only (-1: every family and the stages)
  hipEvent_t ev[NSTAGE + 1];
  double stage_ms[NSTAGE];
  // per-kernel timing: event pairs of this step, accumulated after it
  hipEvent_t kev[2 * MAXEV];
  int kid[MAXEV];
  int nkev;
  double kern_ms[KT_N];
  long kern_n[KT_N];
};

const char *cress_error_string(int code) {
  switch (code) {
    case CRESS_OK: return "ok";
    case CRESS_EINVAL: return "invalid argument";
    case CRESS_ENODEV: return "no usable GPU";
    case CRESS_EHIP: return "HIP runtime error";
    case CRESS_ENOMEM: return "device out of memory";
    case CRESS_ERANGE: return "field id or level out of range";
    case CRESS_EUNSUPPORTED: return "option outside the implemented plain path";
    case CRESS_ETOOBIG: return "grid larger than the fixed scratch buffer";
  }
  return "unknown error";
}

int cress_device_count(int *n) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return c > 0 ? CRESS_OK : CRESS_ENODEV;
}

static int dalloc(cress_ctx *x, float **p, size_t nfloat) {
  void *q = nullptr;
  if (hipMalloc(&q, nfloat * sizeof(float)) != hipSuccess) return CRESS_ENOMEM;
  if (hipMemset(q, 0, nfloat * sizeof(float)) != hipSuccess) return CRESS_EHIP;
  x->allocs.push_back(q);
  *p = (float *)q;
  return CRESS_OK;
}

static int check_config(const cress_config *c) {
  if (!c || c->ni < 7 || c->nj < 7 || c->nk < 7) return CRESS_EINVAL;
  if (c->nsstp < 1 || !(c->dx > 0) || !(c->dy > 0) || !(c->dz > 0)) return CRESS_EINVAL;
  // kernels address within a (row-padded) field by 32-bit byte offsets
  if (make_grid_padded(c->ni, c->nj, c->nk).n() * 4L >= (1L << 32)) return CRESS_EUNSUPPORTED;
  const int periodic = c->wbc == -1 && c->ebc == -1 && c->sbc == -1 && c->nbc == -1;
  const int open = c->wbc == 7 && c->ebc == 7 && c->sbc == 7 && c->nbc == 7;
  if (!periodic && !open) return CRESS_EUNSUPPORTED;
  if (open && !(c->gwave >= 0.0 && c->gwave < 1.0e6)) return CRESS_EINVAL;
  if ((c->bbc != 2 && c->bbc != 3) || (c->tbc != 2 && c->tbc != 3)) return CRESS_EUNSUPPORTED;
  if (c->advopt < 1 || c->advopt > 3 || c->impopt != 1 || c->gwmopt != 0) return CRESS_EUNSUPPORTED;
  if (c->divopt < 0 || c->divopt > 1 || c->buyopt < 0 || c->buyopt > 1) return CRESS_EUNSUPPORTED;
  if ((c->cphopt != 0 && c->cphopt != 2) || c->haiopt != 0) return CRESS_EUNSUPPORTED;
  if (c->vspopt != 0 && c->vspopt != 2) return CRESS_EUNSUPPORTED;
  if (c->smtopt != 0 && c->smtopt != 2 && c->smtopt != 3) return CRESS_EUNSUPPORTED;
  if (c->smtopt && !(c->smhcoe >= 0.0 && c->smvcoe >= 0.0 && c->smhcoe < 1.0e6 && c->smvcoe < 1.0e6))
    return CRESS_EINVAL;
  return CRESS_OK;
}

// Derived binary32 constants, folded in the same order as the reference.
static void fold_consts(const cress_config *cf, Consts *c) {
  const float dx = (float)cf->dx, dy = (float)cf->dy, dz = (float)cf->dz;
  const float oned24 = 1.0f / 24.0f;
  c->dxiv = 1.0f / dx;
  c->dyiv = 1.0f / dy;
  c->dziv = 1.0f / dz;
  c->dxv05n = -0.5f * c->dxiv;
  c->dyv05n = -0.5f * c->dyiv;
  c->dzv05n = -0.5f * c->dziv;
  c->dxv25n = -0.25f * c->dxiv;
  c->dyv25n = -0.25f * c->dyiv;
  c->dzv25n = -0.25f * c->dziv;
  c->dxv24 = oned24 * c->dxiv;
  c->dyv24 = oned24 * c->dyiv;
  c->dzv24 = oned24 * c->dziv;
  c->dxv48 = 0.5f * c->dxv24;
  c->dyv48 = 0.5f * c->dyv24;
  c->dzv48 = 0.5f * c->dzv24;
  c->fourd3 = 4.0f / 3.0f;
  c->g05 = 0.5f * 9.8f;
  c->g05n = -0.5f * 9.8f;
  c->weicoe = (float)cf->weicoe;
  c->weic1m = 1.0f - c->weicoe;
  c->filcoe = (float)cf->filcoe;
  c->m1fc2 = 1.0f - 2.0f * c->filcoe;
  c->dz = dz;
  c->thresq = (float)cf->thresq;
  c->smh = (float)cf->smhcoe;
  c->smv = (float)cf->smvcoe;
}

int cress_ctx_create(const cress_config *cfg, cress_ctx **out) {
  if (!out) return CRESS_EINVAL;
  *out = nullptr;
  int rc = check_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  cress_ctx *x = new cress_ctx();
  x->cfg = *cfg;
  x->g = make_grid_padded(cfg->ni, cfg->nj, cfg->nk);
  x->n = x->g.n();
  fold_consts(cfg, &x->c);
  Opts &o = x->o;
  o.advopt = cfg->advopt;
  o.divopt = cfg->divopt;
  o.buyopt = cfg->buyopt;
  o.bbc = cfg->bbc;
  o.tbc = cfg->tbc;
  o.sthopt = cfg->sthopt;
  o.coropt = cfg->coropt;
  o.cphopt = cfg->cphopt;
  o.iw = cfg->wbc == -1 ? 0 : 1;
  o.ie = cfg->ebc == -1 ? 0 : 1;
  o.js = cfg->sbc == -1 ? 0 : 1;
  o.jn = cfg->nbc == -1 ? 0 : 1;
  const int vs = cfg->vspopt == 2;
  o.vsp_u = vs && cfg->vspvar[0] == 'o';
  o.vsp_v = vs && cfg->vspvar[1] == 'o';
  o.vsp_w = vs && cfg->vspvar[2] == 'o';
  o.vsp_p = vs && cfg->vspvar[3] == 'o';
  o.vsp_pt = vs && cfg->vspvar[4] == 'o';
  o.vsp_qv = vs && cfg->vspvar[5] == 'o';
  o.vsp_q = vs && cfg->vspvar[6] == 'o';
  o.ksp0 = cfg->ksp0[1];
  o.smtopt = cfg->smtopt;
  x->tune = tune_default();
  x->setupcol = nullptr;
  if (hipStreamCreate(&x->stream) != hipSuccess) {
    delete x;
    return CRESS_EHIP;
  }
  const long n = x->n;
  const long n2 = (long)(cfg->ni + 2) * (cfg->nj + 2);
#define AL(p, cnt) if ((rc = dalloc(x, &(p), (size_t)(cnt))) != 0) goto fail
  for (int f = 0; f < CRESS_NBASE; f++) AL(x->base[f], n);
  AL(x->tab, (long)CRESS_NBASE * cfg->nk);
  x->tabh.assign((size_t)CRESS_NBASE * cfg->nk, 0.0f);
  AL(x->accoef, 3L * 3 * (cfg->nk + 1));
  {
    float *dp = nullptr;
    AL(dp, 2L * ((size_t)kDiagBlocks * DIAG_N + DIAG_N + cfg->nk + 1));
    x->diag = (double *)dp;
  }
  for (int s = 0; s < 3; s++) x->accoef_ok[s] = 0;
  x->g.tab = x->tab;
  x->g.umask = 0u;
  AL(x->fc, 2 * n2);
  AL(x->rbct, 2L * cfg->ni * cfg->nj * cfg->nk);
  for (int l = 0; l < 3; l++) {
    AL(x->u[l], n);
    AL(x->v[l], n);
    AL(x->w[l], n);
    AL(x->pp[l], n);
    AL(x->ptp[l], n);
    AL(x->q[l], 6 * n);
    AL(x->nci[l], n);
  }
  AL(x->prwtr, 4 * n2);
  AL(x->price, 6 * n2);
  AL(x->ufrc, n);
  AL(x->vfrc, n);
  AL(x->wfrc, n);
  AL(x->pfrc, n);
  AL(x->ptfrc, n);
  AL(x->qvfrc, n);
  AL(x->wc, n);
  AL(x->ptil, n);
  AL(x->ualt, n);
  AL(x->valt, n);
  AL(x->rstxu, n);
  AL(x->rstxv, n);
  AL(x->rstxwc, n);
  if (cfg->cphopt == 2) {
    for (int f = 0; f < MicroScratch::NF; f++) AL(x->ms.f[f], n);
    AL(x->sfrc, 6 * n);
    AL(x->ms.col, 64);
    rc = micro_init_tables();
    if (rc) goto fail;
  }
#undef AL
  for (int s = 0; s <= NSTAGE; s++) hipEventCreate(&x->ev[s]);
  for (int s = 0; s < 2 * MAXEV; s++) hipEventCreate(&x->kev[s]);
  x->rot = 0;
  *out = x;
  return CRESS_OK;
fail:
  cress_ctx_destroy(x);
  return rc;
}

int cress_ctx_destroy(cress_ctx *x) {
  if (!x) return CRESS_OK;
  hipStreamSynchronize(x->stream);
  for (void *p : x->allocs) hipFree(p);
  for (int s = 0; s <= NSTAGE; s++)
    if (x->ev[s]) hipEventDestroy(x->ev[s]);
  for (int s = 0; s < 2 * MAXEV; s++)
    if (x->kev[s]) hipEventDestroy(x->kev[s]);
  hipStreamDestroy(x->stream);
  delete x;
  return CRESS_OK;
}

// Device address of a field; *count: its float count in the host layout;
// *pad3d: a 3-D field stored with padded rows on the device.
static float *field_ptr(cress_ctx *x, int f, int level, size_t *count, bool *pad3d = nullptr) {
  const long n = x->n;
  if (pad3d) *pad3d = (f >= 0 && f < CRESS_NBASE) || (f >= CRESS_U && f < CRESS_NPROG_END);
  const long n2 = (long)(x->cfg.ni + 2) * (x->cfg.nj + 2);
  *count = (size_t)(x->cfg.ni + 2) * (size_t)(x->cfg.nj + 2) * (size_t)x->cfg.nk;
  if (f >= 0 && f < CRESS_NBASE) return level == 0 ? x->base[f] : nullptr;
  if (f == CRESS_FC) {
    *count = 2 * n2;
    return level == 0 ? x->fc : nullptr;
  }
  if (f == CRESS_RBCT) {
    *count = 2L * x->cfg.ni * x->cfg.nj * x->cfg.nk;
    return level == 0 ? x->rbct : nullptr;
  }
  if (f == CRESS_PRWTR) {
    *count = 4 * n2;
    return level == 0 ? x->prwtr : nullptr;
  }
  if (f == CRESS_PRICE) {
    *count = 6 * n2;
    return level == 0 ? x->price : nullptr;
  }
  if (f >= CRESS_U && f < CRESS_NPROG_END) {
    if (level < 0 || level > 2) return nullptr;
    const int s = (x->rot + level) % 3;
    switch (f) {
      case CRESS_U: return x->u[s];
      case CRESS_V: return x->v[s];
      case CRESS_W: return x->w[s];
      case CRESS_PP: return x->pp[s];
      case CRESS_PTP: return x->ptp[s];
      case CRESS_NCI: return x->nci[s];
      default: return x->q[s] + (long)(f - CRESS_QV) * n;
    }
  }
  return nullptr;
}

int cress_field_size(const cress_ctx *x, int f, size_t *nfloat) {
  size_t c = 0;
  if (!x || !nfloat) return CRESS_EINVAL;
  if (!field_ptr(const_cast<cress_ctx *>(x), f, 0, &c)) return CRESS_ERANGE;
  *nfloat = c;
  return CRESS_OK;
}

// The column kernel's Thomas coefficients for one dts when rst8w, rcsq, rst,
// rbr and jcb (and, for the kernel's other reads, jcb8u, jcb8v, jcb8w) are
// horizontally uniform (AcCol::coef), formed on the host with
// the kernel's own binary32 operations in its order (stepwi.f90:391-416,
// gaussel.f90:189-236); the build keeps FMA contraction off on the host too,
// so these are the bits each column would compute.
static bool ac_uniform(const cress_ctx *x) {
  const unsigned need = (1u << CRESS_RST8W) | (1u << CRESS_RCSQ) | (1u << CRESS_RST) | (1u << CRESS_RBR) |
                        (1u << CRESS_JCB) | (1u << CRESS_JCB8U) | (1u << CRESS_JCB8V) | (1u << CRESS_JCB8W);
  return (x->g.umask & need) == need;
}

static void ac_coef_host(const cress_ctx *x, float dts, float *coef) {
  const Consts &c = x->c;
  const int nk = x->g.nk;
  auto tab = [&](int id, int k) { return x->tabh[(size_t)id * nk + k - 1]; };
  const float sbsqzi = dts * dts * c.weicoe * c.weicoe * c.dziv;
  const float sbsqg5 = c.g05 * dts * dts * c.weicoe * c.weicoe;
  float fwc_km1 = 0.0f, wcc_km1 = 0.0f, rcsq_km1 = 0.0f, rst_km1 = 0.0f, ssn_km1 = 0.0f;
  for (int k = 0; k < 3 * (nk + 1); k++) coef[k] = 0.0f;
  for (int k = 2; k <= nk - 2; k++) {
    const float rcsq = tab(CRESS_RCSQ, k), rst = tab(CRESS_RST, k);
    const float jcb = tab(CRESS_JCB, k), rbr = tab(CRESS_RBR, k);
    const float t1 = c.g05 * rbr / rcsq;
    const float t2 = c.dziv / jcb;
    const float fwc = t1 + t2, wcc = t1 - t2;
    if (k >= 3) {
      const float rst8w = tab(CRESS_RST8W, k);
      float rr, ss, tt;
      if (x->o.buyopt == 0) {
        const float aa = sbsqzi / rst8w;
        const float mm = aa * rcsq_km1;
        const float nn = aa * rcsq;
        rr = -mm * fwc_km1;
        ss = 1.0f + (nn * fwc - mm * wcc_km1);
        tt = nn * wcc;
      } else {
        const float rstiv = 1.0f / rst8w;
        const float aa = rstiv * sbsqzi;
        const float bb = rstiv * sbsqg5;
        const float mm = bb * rst_km1 - aa * rcsq_km1;
        const float nn = bb * rst + aa * rcsq;
        rr = mm * fwc_km1;
        ss = 1.0f + (mm * wcc_km1 + nn * fwc);
        tt = nn * wcc;
      }
      const float r = k == 3 ? 1.0f / ss : 1.0f / (ss - rr * ssn_km1);
      const float ssn = tt * r;
      coef[3 * k] = rr;
      coef[3 * k + 1] = r;
      coef[3 * k + 2] = ssn;
      ssn_km1 = ssn;
    }
    fwc_km1 = fwc;
    wcc_km1 = wcc;
    rcsq_km1 = rcsq;
    rst_km1 = rst;
  }
}

static int upload_accoef(cress_ctx *x, int slot, float dts) {
  const int nk = x->g.nk;
  std::vector<float> h(3 * (size_t)(nk + 1));
  ac_coef_host(x, dts, h.data());
  HIPCHK(hipMemcpyAsync(x->accoef + (size_t)slot * 3 * (nk + 1), h.data(), h.size() * sizeof(float),
                        hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  x->accoef_dts[slot] = dts;
  x->accoef_ok[slot] = 1;
  return CRESS_OK;
}

// After any change of the base fields: the tables of the large step's two
// substep lengths (the cold start's dtsml/2 and dtsml), or none.
static int refresh_accoef(cress_ctx *x) {
  for (int s = 0; s < 3; s++) x->accoef_ok[s] = 0;
  if (!ac_uniform(x)) return CRESS_OK;
  int rc = upload_accoef(x, 0, 0.5f * (float)x->cfg.dtsml);
  if (!rc) rc = upload_accoef(x, 1, (float)x->cfg.dtsml);
  return rc;
}

// The table for dts, or null (non-uniform metrics: the kernel forms the
// coefficients per column).  A dts other than the large step's two (the
// per-substep entry point) is formed and uploaded on demand, synchronously.
static const float *accoef_for(cress_ctx *x, float dts) {
  if (!ac_uniform(x)) return nullptr;
  const size_t stride = 3 * (size_t)(x->g.nk + 1);
  for (int s = 0; s < 3; s++)
    if (x->accoef_ok[s] && x->accoef_dts[s] == dts) return x->accoef + s * stride;
  if (upload_accoef(x, 2, dts) != CRESS_OK) return nullptr;
  return x->accoef + 2 * stride;
}

// A base field whose columns all hold the same bits, level by level: its
// per-level values go to the table and the kernels read them from there.
static int note_uniform(cress_ctx *x, int f, const float *host) {
  const long nxy = (long)(x->cfg.ni + 2) * (x->cfg.nj + 2);   // host layout
  const int nk = x->g.nk;
  bool uni = true;
  for (int k = 0; k < nk && uni; k++) {
    const float *pl = host + (long)k * nxy;
    for (long p = 1; p < nxy; p++)
      if (memcmp(pl + p, pl, sizeof(float)) != 0) {
        uni = false;
        break;
      }
  }
  if (uni) {
    std::vector<float> col(nk);
    for (int k = 0; k < nk; k++) col[k] = host[(long)k * nxy];
    HIPCHK(hipMemcpyAsync(x->tab + (long)f * nk, col.data(), nk * sizeof(float), hipMemcpyHostToDevice,
                          x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int k = 0; k < nk; k++) x->tabh[(size_t)f * nk + k] = col[k];
    x->g.umask |= 1u << f;
  } else {
    x->g.umask &= ~(1u << f);
  }
  return refresh_accoef(x);
}

// Host rows of ni+2 floats <-> device rows of g.nx floats.
static int copy_field(cress_ctx *x, float *d, const float *h, bool up) {
  const size_t hrow = (size_t)(x->cfg.ni + 2) * sizeof(float), drow = (size_t)x->g.nx * sizeof(float);
  const size_t rows = (size_t)(x->cfg.nj + 2) * (size_t)x->cfg.nk;
  if (up)
    HIPCHK(hipMemcpy2DAsync(d, drow, h, hrow, hrow, rows, hipMemcpyHostToDevice, x->stream));
  else
    HIPCHK(hipMemcpy2DAsync((void *)h, hrow, d, drow, hrow, rows, hipMemcpyDeviceToHost, x->stream));
  return CRESS_OK;
}

int cress_upload(cress_ctx *x, int f, int level, const float *host) {
  size_t c = 0;
  bool p3 = false;
  if (!x || !host) return CRESS_EINVAL;
  float *d = field_ptr(x, f, level, &c, &p3);
  if (!d) return CRESS_ERANGE;
  if (p3) {
    int rc = copy_field(x, d, host, true);
    if (rc) return rc;
  } else {
    HIPCHK(hipMemcpyAsync(d, host, c * sizeof(float), hipMemcpyHostToDevice, x->stream));
  }
  HIPCHK(hipStreamSynchronize(x->stream));
  if (f >= 0 && f < CRESS_NBASE) return note_uniform(x, f, host);
  return CRESS_OK;
}

int cress_download(cress_ctx *x, int f, int level, float *host) {
  size_t c = 0;
  bool p3 = false;
  if (!x || !host) return CRESS_EINVAL;
  float *d = field_ptr(x, f, level, &c, &p3);
  if (!d) return CRESS_ERANGE;
  if (p3) {
    int rc = copy_field(x, d, host, false);
    if (rc) return rc;
  } else {
    HIPCHK(hipMemcpyAsync(host, d, c * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  }
  HIPCHK(hipStreamSynchronize(x->stream));
  return CRESS_OK;
}

int cress_field_device_ptr(cress_ctx *x, int f, int level, void **dptr) {
  size_t c = 0;
  if (!x || !dptr) return CRESS_EINVAL;
  float *d = field_ptr(x, f, level, &c);
  if (!d) return CRESS_ERANGE;
  // the caller may write the field: read it as a 3-D field from now on
  if (f >= 0 && f < CRESS_NBASE) {
    x->g.umask &= ~(1u << f);
    int rc = refresh_accoef(x);
    if (rc) return rc;
  }
  *dptr = d;
  return CRESS_OK;
}

int cress_field_pitch(const cress_ctx *x, size_t *row, size_t *plane) {
  if (!x || !row || !plane) return CRESS_EINVAL;
  *row = (size_t)x->g.nx;
  *plane = (size_t)x->g.nxy;
  return CRESS_OK;
}

int cress_sync(cress_ctx *x) {
  if (!x) return CRESS_EINVAL;
  HIPCHK(hipStreamSynchronize(x->stream));
  return CRESS_OK;
}

int cress_set_timing(cress_ctx *x, int on) {
  if (!x || on < 0 || on >= 2 + KT_N) return CRESS_EINVAL;
  x->timing = on != 0;
  x->tkern = on >= 2 ? on - 2 : -1;
  return CRESS_OK;
}

int cress_stage_times(cress_ctx *x, double *ms, int nstage) {
  if (!x || !ms) return CRESS_EINVAL;
  for (int s = 0; s < nstage && s < NSTAGE; s++) ms[s] = x->stage_ms[s];
  return CRESS_OK;
}

int cress_kernel_count(void) { return KT_N; }

const char *cress_kernel_name(int id) { return id >= 0 && id < KT_N ? kKernelNames[id] : ""; }

int cress_kernel_times(cress_ctx *x, double *total_ms, long *launches, int n) {
  if (!x || !total_ms || !launches) return CRESS_EINVAL;
  for (int k = 0; k < n && k < KT_N; k++) {
    total_ms[k] = x->kern_ms[k];
    launches[k] = x->kern_n[k];
  }
  return CRESS_OK;
}

int cress_reset_timing(cress_ctx *x) {
  if (!x) return CRESS_EINVAL;
  for (int k = 0; k < KT_N; k++) {
    x->kern_ms[k] = 0.0;
    x->kern_n[k] = 0;
  }
  for (int s = 0; s < NSTAGE; s++) x->stage_ms[s] = 0.0;
  return CRESS_OK;
}

// Launch with optional event bracketing.
#define LAUNCH(x, id, kern, grid, block, lds, ...)                                     \
  do {                                                                                 \
    const bool t_ = (x)->timing && (x)->nkev < MAXEV && ((x)->tkern < 0 || (x)->tkern == (id)); \
    if (t_) hipEventRecord((x)->kev[2 * (x)->nkev], (x)->stream);                      \
    hipLaunchKernelGGL(kern, grid, block, lds, (x)->stream, __VA_ARGS__);              \
    if (t_) {                                                                          \
      hipEventRecord((x)->kev[2 * (x)->nkev + 1], (x)->stream);                        \
      (x)->kid[(x)->nkev++] = (id);                                                    \
    }                                                                                  \
  } while (0)

static void flush_kernel_times(cress_ctx *x) {
  if (!x->nkev) return;
  hipEventSynchronize(x->kev[2 * x->nkev - 1]);
  for (int e = 0; e < x->nkev; e++) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, x->kev[2 * e], x->kev[2 * e + 1]);
    x->kern_ms[x->kid[e]] += ms;
    x->kern_n[x->kid[e]] += 1;
  }
  x->nkev = 0;
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
static inline dim3 grid2(int ni, int nj, int bx, int by, int nz) {
  return dim3((unsigned)((ni + bx - 1) / bx), (unsigned)((nj + by - 1) / by), (unsigned)nz);
}

static void fill_spec(FillSpec *s, float *p, int iws, int iwr, int ies, int ier, int jss, int jsr,
                      int jns, int jnr, int kmax) {
  s->ptr = p;
  s->iwsnd = iws;
  s->iwrcv = iwr;
  s->iesnd = ies;
  s->iercv = ier;
  s->jssnd = jss;
  s->jsrcv = jsr;
  s->jnsnd = jns;
  s->jnrcv = jnr;
  s->kmax = kmax;
}

static void run_fill(cress_ctx *x, const FillSet &fs) {
  if (fs.n == 0) return;
  const Grid &g = x->g;
  int kmax = 0;
  for (int f = 0; f < fs.n; f++) kmax = fs.f[f].kmax > kmax ? fs.f[f].kmax : kmax;
  LAUNCH(x, KT_FILL, k_bcycle_x, dim3((g.nj + 2 + 63) / 64, kmax, fs.n), dim3(64), 0, g, fs);
  LAUNCH(x, KT_FILL, k_bcycle_y, dim3((g.ni + 2 + 63) / 64, kmax, fs.n), dim3(64), 0, g, fs);
}

static void run_vbc(cress_ctx *x, const VbcSet &vs) {
  const Grid &g = x->g;
  LAUNCH(x, KT_VBC, k_vbc, grid2(g.ni, g.nj, 64, 4, vs.n), dim3(64, 4), 0, g, x->o.bbc, vs);
}

static inline void stamp(cress_ctx *x, int s) {
  if (x->timing && x->tkern < 0) hipEventRecord(x->ev[s], x->stream);
}

// Phase speeds of the open-side fills for the coming step (phvuvw.f90:334-346
// and the option-7 branches; phvs.f90:270-290), from the previous step's dtb
// and dts, in binary32 in the reference's order: gdxdt = gwave*dtb*dxiv,
// clipped to [-1,1]; u, v, w and p' scale it by dts/dtb, the large-step
// scalars use gamma = 2c/(1-+c).  Zero before the first step of a cold start
// (allocslv.f90:2015); a context whose first step is not the model's first
// takes the previous step to have been a full one (dtbig, dtsml), which is
// what a restart file written after it holds.
static void obc_speeds(cress_ctx *x, int first, ObcAc *ac, ObcLf *lf) {
  *ac = ObcAc{0.0f, 0.0f, 0.0f, 0.0f};
  *lf = ObcLf{0.0f, 0.0f, 0.0f, 0.0f};
  if (!x->obc_have) {
    if (first) return;
    x->obc_dtb = (float)x->cfg.dtbig;
    x->obc_dts = (float)x->cfg.dtsml;
    x->obc_have = 1;
  }
  const float dtb = x->obc_dtb, dts = x->obc_dts;
  const float gwave = (float)x->cfg.gwave;
  const float gx = gwave * dtb * x->c.dxiv, gy = gwave * dtb * x->c.dyiv;
  const float bw = fmaxf(-gx, -1.0f), be = fminf(gx, 1.0f);
  const float bs = fmaxf(-gy, -1.0f), bn = fminf(gy, 1.0f);
  const float r = dts / dtb;
  ac->cw = bw * r;
  ac->ce = be * r;
  ac->cs = bs * r;
  ac->cn = bn * r;
  lf->gw = 2.0f * bw / (1.0f - bw);
  lf->ge = 2.0f * be / (1.0f + be);
  lf->gs = 2.0f * bs / (1.0f - bs);
  lf->gn = 2.0f * bn / (1.0f + bn);
}

static void corner_spec(CornerSet *cs, float *p, int isw, int ise, int jss, int jsn) {
  const int f = cs->n++;
  cs->ptr[f] = p;
  cs->isw[f] = isw;
  cs->ise[f] = ise;
  cs->jss[f] = jss;
  cs->jsn[f] = jsn;
}

static void run_corners(cress_ctx *x, const CornerSet &cs) {
  if (cs.n == 0) return;
  LAUNCH(x, KT_OBC, k_obc_corners, dim3((x->g.nk + 63) / 64, cs.n), dim3(64), 0, x->g, cs);
}

// ---------------------------------------------------------------------------
// One large step.
// ---------------------------------------------------------------------------
// One acoustic substep (s_hevi, hevi.f90:339-427, with stepuv/stepwi/
// steppi's fills) on the future level: k_ac_ptuv, the u/v fills, the column
// kernel, the w/p' fills and the ring columns.  oac: the open sides' phase
// speeds (null on periodic sides).  uf and vf leave in the other buffer of
// their pair.
// Arguments of the fused ptil/stepuv kernel and of the column kernel for
// one substep of length dts on the context's current future level (uf, vf
// read from u[F], v[F]; the ptil/stepuv kernel writes the other buffer of
// each pair).
static AcPtUv ptuv_args(cress_ctx *x, float dts) {
  const int F = (x->rot + 2) % 3;
  float **B = x->base;
  const float dz = (float)x->cfg.dz;
  AcPtUv pu;
  pu.dts = dts;
  pu.divcv = 0.05f / dts * dz * dz;
  pu.jcb = B[CRESS_JCB];
  pu.rst8u = B[CRESS_RST8U];
  pu.rst8v = B[CRESS_RST8V];
  pu.rst8w = B[CRESS_RST8W];
  pu.wc = x->wc;
  pu.ppf = x->pp[F];
  pu.ufrc = x->ufrc;
  pu.vfrc = x->vfrc;
  pu.ptil = x->ptil;
  pu.uf = x->u[F];
  pu.vf = x->v[F];
  pu.ufn = x->ualt;
  pu.vfn = x->valt;
  return pu;
}

static AcCol col_args(cress_ctx *x, float dts) {
  const int F = (x->rot + 2) % 3;
  float **B = x->base;
  AcCol ac;
  ac.dts = dts;
  ac.ptil = x->ptil;
  ac.pfrc = x->pfrc;
  ac.wfrc = x->wfrc;
  ac.jcb = B[CRESS_JCB];
  ac.jcb8u = B[CRESS_JCB8U];
  ac.jcb8v = B[CRESS_JCB8V];
  ac.jcb8w = B[CRESS_JCB8W];
  ac.rcsq = B[CRESS_RCSQ];
  ac.rst = B[CRESS_RST];
  ac.rbr = B[CRESS_RBR];
  ac.rst8w = B[CRESS_RST8W];
  ac.wf = x->w[F];
  ac.wc = x->wc;
  ac.ppf = x->pp[F];
  ac.uf = x->u[F];
  ac.vf = x->v[F];
  ac.coef = accoef_for(x, dts);
  return ac;
}

static void launch_ptuv(cress_ctx *x, const LaunchTune &tn, const AcPtUv &pu) {
  const Grid g = x->g;
  const Tiles tpu = tiles(g.ni, g.nj - 1, 64, tn.ptuv_by, g.nk - 1, tn.ptuv_order);
  const auto kpu = ptuv_kernel(g, tn.ptuv_slots);
  LAUNCH(x, KT_AC_PTUV, kpu, flat(tpu), dim3(64, tn.ptuv_by), 0, tpu, g, x->o, x->c, pu);
}

static void launch_col(cress_ctx *x, const LaunchTune &tn, const AcCol &ac) {
  const Grid g = x->g;
  const Tiles tcol = tiles(g.ni - 1, g.nj - 3, 64, tn.col_by, 1, 1);
  const auto kcol = ac.coef ? k_ac_col<true> : k_ac_col<false>;
  LAUNCH(x, KT_AC_COL, kcol, flat(tcol), dim3(64, tn.col_by), tn.col_lds, tcol, g, x->o, x->c, ac);
}

static void ac_substep(cress_ctx *x, float dts, const ObcAc *oacp) {
  const Grid g = x->g;
  const Opts o = x->o;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const int F = (x->rot + 2) % 3;
  float **B = x->base;
  const int open = oacp != nullptr;
  const ObcAc oac = open ? *oacp : ObcAc{0.0f, 0.0f, 0.0f, 0.0f};
  FillSet fwp;
  fwp.n = 2;
  fill_spec(&fwp.f[0], x->w[F], 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk);
  fill_spec(&fwp.f[1], x->pp[F], 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk);
  const int nring = 2 * (ni - 1) + 2 * (nj - 3);
  {
    launch_ptuv(x, x->tune, ptuv_args(x, dts));
    std::swap(x->u[F], x->ualt);
    std::swap(x->v[F], x->valt);
    if (open) {
      // rbcu/rbcv from the previous substep's u, v (stepuv.f90:266-280),
      // rbcw and rbcss on w, p' before the column kernel updates the
      // interior next to them (stepwi.f90:286-296, steppi), then the corner
      // means of stepuv, stepwi and steppi (bc4news)
      ObcAcArgs oa;
      oa.uold = x->ualt;
      oa.vold = x->valt;
      oa.unew = x->u[F];
      oa.vnew = x->v[F];
      oa.wf = x->w[F];
      oa.ppf = x->pp[F];
      LAUNCH(x, KT_OBC, k_obc_ac_edges, dim3((2 * (nj + 2) + 2 * (ni + 2) + 63) / 64, nk), dim3(64), 0, g,
             oac, oa);
      CornerSet cs;
      cs.n = 0;
      corner_spec(&cs, x->u[F], 1, ni, 1, nj - 1);
      corner_spec(&cs, x->v[F], 1, ni - 1, 1, nj);
      corner_spec(&cs, x->w[F], 1, ni - 1, 1, nj - 1);
      corner_spec(&cs, x->pp[F], 1, ni - 1, 1, nj - 1);
      run_corners(x, cs);
    }
    FillSet fuv;
    fuv.n = open ? 0 : 2;
    fill_spec(&fuv.f[0], x->u[F], 3, 1, ni - 2, ni, 2, 1, nj - 2, nj - 1, nk);
    fill_spec(&fuv.f[1], x->v[F], 2, 1, ni - 2, ni - 1, 3, 1, nj - 2, nj, nk);
    VbcSet vuv;
    vuv.n = 2;
    vuv.kind[0] = 1;
    vuv.ptr[0] = x->u[F];
    vuv.kind[1] = 2;
    vuv.ptr[1] = x->v[F];
    run_fill(x, fuv);
    run_vbc(x, vuv);
    launch_col(x, x->tune, col_args(x, dts));
    if (!open) run_fill(x, fwp);
    LAUNCH(x, KT_AC_RING, k_ac_ring, dim3((nring + 63) / 64, nk), dim3(64), 0, g, o, B[CRESS_JCB8W], x->w[F],
           x->wc, x->pp[F]);
  }
}

// s_cloudphy for cphopt=2 on the future level (cloudphy.f90:277-338 ->
// coldblk.f90:321-481): k_micro_src (with the scalar leapfrog fused unless
// the future level is given) and k_micro_fall.
static void micro_stage(cress_ctx *x, float dtb, const ObcLf &olf, int open, int given) {
  const Grid g = x->g;
  const Consts c = x->c;
  const cress_config &cf = x->cfg;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const long n = x->n;
  const int P = (x->rot + 0) % 3, C = (x->rot + 1) % 3, F = (x->rot + 2) % 3;
  float **B = x->base;
  const dim3 b2(64, 4);
  {
    // leapfrog: the microphysics step is 2*dtb (cloudphy.f90:281-285)
    const float dtbs = 2.0f * dtb;
    const MicroK mk = micro_fold(dtbs, cf.thresq, (float)cf.dz);
    MicroArgs ma;
    ma.jcb = B[CRESS_JCB];
    ma.pbr = B[CRESS_PBR];
    ma.ptbr = B[CRESS_PTBR];
    ma.rbr = B[CRESS_RBR];
    ma.ppp = x->pp[P];
    ma.ptpp = x->ptp[P];
    ma.qp = x->q[P];
    ma.qpw = x->q[P] + n;
    ma.qpi = x->q[P] + 3 * n;
    ma.ncip = x->nci[P];
    ma.ptpf = x->ptp[F];
    ma.rst = B[CRESS_RST];
    ma.ptfrc = x->ptfrc;
    ma.qvfrc = x->qvfrc;
    ma.sfrc = x->sfrc;
    ma.qf = x->q[F];
    ma.qfw = x->q[F] + n;
    ma.qfi = x->q[F] + 3 * n;
    ma.ncif = x->nci[F];
    ma.urq = x->ms.f[0];
    ma.usq = x->ms.f[1];
    ma.ugq = x->ms.f[2];
    ma.dtpmin = (unsigned *)x->ms.col;
    ma.qn = n;
    ma.ptpc = x->ptp[C];
    ma.qc = x->q[C];
    ma.ncic = x->nci[C];
    ma.lf = olf;
    ma.open = open;
    ma.given = given;
    ma.mode = x->mlp ? x->mlp_mode : 0;
    ma.cap = MlpCapture{nullptr, nullptr, nullptr, 0};
    ma.mflag = nullptr;
    ma.mout = nullptr;
    if (ma.mode == 1) {
      const cress_mlp_impl *m = (const cress_mlp_impl *)x->mlp;
      ma.cap = MlpCapture{m->din, m->dtgt, m->dcount, m->cap};
    } else if (ma.mode == 2) {
      const cress_mlp_impl *m = (const cress_mlp_impl *)x->mlp;
      MlpSelectArgs sa;
      sa.pbr = B[CRESS_PBR];
      sa.ptbr = B[CRESS_PTBR];
      sa.rbr = B[CRESS_RBR];
      sa.ppp = x->pp[P];
      sa.ptpp = x->ptp[P];
      sa.qp = ma.qp;
      sa.qpw = ma.qpw;
      sa.qpi = ma.qpi;
      sa.qn = n;
      sa.thresq = (float)cf.thresq;
      sa.s = x->ml;
      sa.norm = m->norm;
      LAUNCH(x, KT_MLP_SELECT, k_u32_set, dim3(1), dim3(64), 0, x->ml.count, 0u);
      const Tiles tsel = tiles(ni, nj - 1, 64, 4, nk - 1);
      LAUNCH(x, KT_MLP_SELECT, k_mlp_select, flat(tsel), b2, 0, tsel, g, sa);
      const MlpNet net{m->w, m->p, m->norm};
      LAUNCH(x, KT_MLP_INFER, k_mlp_infer, dim3(256), dim3(MLP_INFER_THREADS), 0, net, x->ml, n);
      ma.mflag = x->ml.flag;
      ma.mout = x->ml.out;
    }
    unsigned seed;
    memcpy(&seed, &dtbs, sizeof seed);
    LAUNCH(x, KT_MICRO_SEED, k_micro_seed, dim3(1), dim3(64), 0, ma.dtpmin, seed);
    LAUNCH(x, KT_MICRO_SRC, k_micro_src, flat(tiles(ni, nj - 1, 64, 4, nk - 1)), b2, 0, tiles(ni, nj - 1, 64, 4, nk - 1), g, mk, ma);
    FallArgs fa;
    fa.rbr = B[CRESS_RBR];
    fa.rst = B[CRESS_RST];
    fa.qpw = ma.qpw;
    fa.qpi = ma.qpi;
    fa.urq = ma.urq;
    fa.usq = ma.usq;
    fa.ugq = ma.ugq;
    fa.qfw = ma.qfw;
    fa.qfi = ma.qfi;
    fa.ncif = x->nci[F];
    fa.prwtr = x->prwtr;
    fa.price = x->price;
    fa.dtpmin = ma.dtpmin;
    fa.qn = n;
    fa.n2 = (long)(ni + 2) * (nj + 2);
    fa.dtb = dtbs;
    fa.dziv = c.dziv;
    fa.thresq = c.thresq;
    const Tiles tfall = tiles(ni - CRESS_FALL_ORG, nj - 1, 64, 1, 1, 1);
    LAUNCH(x, KT_MICRO_FALL, k_micro_fall, flat(tfall), dim3(64, 6), 0, tfall, g, fa);
  }
}

// The forcing kernels (forcedrv) on the present level with launch shapes
// tn; parts selects kernels (all of them in the large step, one family at a
// time in cress_autotune).
enum { FORCE_PREP = 1, FORCE_UV = 2, FORCE_W = 4, FORCE_P = 8, FORCE_S = 16, FORCE_ALL = 31 };
static void forcing_stage(cress_ctx *x, const LaunchTune &tn, unsigned parts) {
  const Grid g = x->g;
  const Opts o = x->o;
  const Consts c = x->c;
  const cress_config &cf = x->cfg;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const long n = x->n;
  const int P = (x->rot + 0) % 3, C = (x->rot + 1) % 3;
  float **B = x->base;
  const dim3 b2(64, 4);
  // ---- forcing (forcedrv) on the present level ----
  // every base/metric field a per-level table: the kernels' table builds
  const bool U = (g.umask & ((1u << CRESS_NBASE) - 1u)) == ((1u << CRESS_NBASE) - 1u);
  const auto kprep = U ? k_prep<true> : k_prep<false>;
  if (parts & FORCE_PREP)
    LAUNCH(x, KT_PREP, kprep, flat(tiles(ni + 2, nj + 2, 64, 4, nk)), b2, 0, tiles(ni + 2, nj + 2, 64, 4, nk), g, o, x->u[C], x->v[C],
                     x->w[C], B[CRESS_RST8U], B[CRESS_RST8V], B[CRESS_RST8W], B[CRESS_JCB8W], x->wc,
                     x->rstxu, x->rstxv, x->rstxwc);
  ForceIn fi;
  fi.u = x->u[C];
  fi.v = x->v[C];
  fi.w = x->w[C];
  fi.up = x->u[P];
  fi.vp = x->v[P];
  fi.wp = x->w[P];
  fi.ptp = x->ptp[C];
  fi.q = x->q[C];
  fi.qvbr = B[CRESS_QVBR];
  fi.ptbr = B[CRESS_PTBR];
  fi.ubr = B[CRESS_UBR];
  fi.vbr = B[CRESS_VBR];
  fi.rst = B[CRESS_RST];
  fi.rst8u = B[CRESS_RST8U];
  fi.rst8v = B[CRESS_RST8V];
  fi.rst8w = B[CRESS_RST8W];
  fi.rstxu = x->rstxu;
  fi.rstxv = x->rstxv;
  fi.rstxwc = x->rstxwc;
  fi.fc = x->fc;
  fi.rbct = x->rbct;
  fi.jcb8u = B[CRESS_JCB8U];
  fi.jcb8v = B[CRESS_JCB8V];
  fi.jcb8w = B[CRESS_JCB8W];
  fi.qn = n;
  // the forcing kernels come in two builds: with the 4th-order smoothing
  // (smtopt 2/3) and without, so that the plain path carries none of its
  // registers
  const bool sm = o.smtopt != 0;
  const dim3 bf(64, tn.ff_by), bs(64, tn.fs_by);
  const Tiles tuv = tiles(ni, nj - 2, 64, tn.ff_by, nk - 3, tn.ff_order);
  const Tiles tw = tiles(ni - 1, nj - 3, 64, tn.ff_by, nk - 2, tn.ff_order);
  const Tiles tp = tiles(ni - 1, nj - 3, 64, tn.ff_by, nk - 3, tn.ff_order);
  const Tiles ts = tiles(ni - 1, nj - 3, 64, tn.fs_by, nk - 3, tn.fs_order);
  const auto kuv = sm ? (U ? k_force_uv<true, true> : k_force_uv<true, false>)
                      : (U ? k_force_uv<false, true> : k_force_uv<false, false>);
  const auto kw = sm ? (U ? k_force_w<true, true> : k_force_w<true, false>)
                     : (U ? k_force_w<false, true> : k_force_w<false, false>);
  const auto kp = U ? k_force_p<true> : k_force_p<false>;
  if (parts & FORCE_UV) LAUNCH(x, KT_FORCE_UV, kuv, flat(tuv), bf, 0, tuv, g, o, c, fi, x->ufrc, x->vfrc);
  if (parts & FORCE_W) LAUNCH(x, KT_FORCE_W, kw, flat(tw), bf, 0, tw, g, o, c, fi, x->wfrc);
  if (parts & FORCE_P)
    LAUNCH(x, KT_FORCE_P, kp, flat(tp), bf, 0, tp, g, o, c, B[CRESS_JCB], B[CRESS_JCB8U], B[CRESS_JCB8V],
         B[CRESS_JCB8W], x->u[C], x->v[C], x->wc, x->pp[C], x->pp[P], x->rbct, x->pfrc);
  ScalarIn si;
  const int nsc = cf.cphopt == 2 ? 8 : 2;
  {
    const float *sC[8] = {x->ptp[C], x->q[C], x->q[C] + n, x->q[C] + 2 * n, x->q[C] + 3 * n,
                          x->q[C] + 4 * n, x->q[C] + 5 * n, x->nci[C]};
    const float *sP[8] = {x->ptp[P], x->q[P], x->q[P] + n, x->q[P] + 2 * n, x->q[P] + 3 * n,
                          x->q[P] + 4 * n, x->q[P] + 5 * n, x->nci[P]};
    // with microphysics the species forcings get their own arrays (the
    // fused steps+microphysics kernel reads them at a cell's periodic image
    // while writing the future level); without, only theta' and qv step
    float *fr[8] = {x->ptfrc, x->qvfrc, x->sfrc, x->sfrc + n, x->sfrc + 2 * n,
                    x->sfrc + 3 * n, x->sfrc + 4 * n, x->sfrc + 5 * n};
    for (int s = 0; s < 8; s++) {
      si.s[s] = sC[s];
      si.sp[s] = sP[s];
      si.frc[s] = fr[s];
    }
  }
  si.rstxu = x->rstxu;
  si.rstxv = x->rstxv;
  si.rstxwc = x->rstxwc;
  si.rst = B[CRESS_RST];
  si.qvbr = B[CRESS_QVBR];
  si.rbct = x->rbct;
  si.ptbr = B[CRESS_PTBR];
  si.rbr = B[CRESS_RBR];
  si.w = x->w[C];
  const auto ks = sm ? (U ? k_force_s<true, true> : k_force_s<true, false>)
                     : (U ? k_force_s<false, true> : k_force_s<false, false>);
  if (parts & FORCE_S) LAUNCH(x, KT_FORCE_S, ks, flat(ts), bs, 0, ts, g, o, c, si, nsc);
}

static int large_step(cress_ctx *x, int first) {
  const Grid g = x->g;
  const Opts o = x->o;
  const Consts c = x->c;
  const cress_config &cf = x->cfg;
  hipStream_t st = x->stream;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const long n = x->n;
  const int P = (x->rot + 0) % 3, C = (x->rot + 1) % 3, F = (x->rot + 2) % 3;
  const float dtb = first ? 0.5f * (float)cf.dtbig : (float)cf.dtbig;
  const float dts = first ? 0.5f * (float)cf.dtsml : (float)cf.dtsml;
  float **B = x->base;
  const dim3 b2(64, 4);
  const int open = cf.wbc == 7;
  ObcAc oac;
  ObcLf olf;
  obc_speeds(x, first, &oac, &olf);

  stamp(x, 0);
  forcing_stage(x, x->tune, FORCE_ALL);
  stamp(x, 1);

  // ---- acoustic substeps (smalldrv / hevi) on the future level ----
  // uf and vf alternate between two buffers each substep (the fused
  // ptil/stepuv kernel reads one, writes the other).  The cells no substep
  // writes (the second halo ring) are never read during the substeps, and
  // after them must hold the past level, as the reference's uf starts as a
  // copy of up: the buffer that ends as the future level starts as that
  // copy -- the future slot itself for an even substep count, the other
  // buffer for an odd one.
  HIPCHK(hipMemcpyAsync(x->u[F], x->u[P], n * sizeof(float), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(x->v[F], x->v[P], n * sizeof(float), hipMemcpyDeviceToDevice, st));
  if (cf.nsstp % 2) {
    HIPCHK(hipMemcpyAsync(x->ualt, x->u[P], n * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(x->valt, x->v[P], n * sizeof(float), hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipMemcpyAsync(x->w[F], x->w[P], n * sizeof(float), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(x->pp[F], x->pp[P], n * sizeof(float), hipMemcpyDeviceToDevice, st));
  const bool U = (g.umask & ((1u << CRESS_NBASE) - 1u)) == ((1u << CRESS_NBASE) - 1u);
  const auto kph = U ? k_phy2cnt<true> : k_phy2cnt<false>;
  LAUNCH(x, KT_PHY2CNT, kph, flat(tiles(ni, nj - 1, 64, 4, nk)), b2, 0, tiles(ni, nj - 1, 64, 4, nk), g,
         o, x->w[F], B[CRESS_JCB8W], x->wc);
  for (int s = 0; s < cf.nsstp; s++) ac_substep(x, dts, open ? &oac : nullptr);
  stamp(x, 2);

  // ---- scalar leapfrog (steps) ----
  // With microphysics the leapfrog and its halo fills are done inside the
  // microphysics source kernel, cell by cell (below).
  const int withq = cf.cphopt == 2;
  if (!withq) {
    StepsIn sti;
    sti.ptpp = x->ptp[P];
    sti.ptfrc = x->ptfrc;
    sti.qp = x->q[P];
    sti.qvfrc = x->qvfrc;
    sti.ncip = x->nci[P];
    sti.ptpf = x->ptp[F];
    sti.qf = x->q[F];
    sti.ncif = x->nci[F];
    sti.qn = n;
    LAUNCH(x, KT_STEPS, k_steps, flat(tiles(ni - 1, nj - 3, 64, 4, nk - 3)), b2, 0,
           tiles(ni - 1, nj - 3, 64, 4, nk - 3), g, 2.0f * dtb, B[CRESS_RST], sti, 0);
    if (open) {
      ObcStepsArgs osa;
      osa.s[0] = x->ptp[C];
      osa.sp[0] = x->ptp[P];
      osa.sf[0] = x->ptp[F];
      osa.s[1] = x->q[C];
      osa.sp[1] = x->q[P];
      osa.sf[1] = x->q[F];
      const int nring = 2 * (ni - 1) + 2 * (nj - 3);
      LAUNCH(x, KT_OBC, k_obc_steps, dim3((nring + 63) / 64, nk - 3, 2), dim3(64), 0, g, olf, osa);
    }
    FillSet fs;
    VbcSet vs;
    fs.n = 2;
    vs.n = fs.n;
    float *fl[2] = {x->ptp[F], x->q[F]};
    for (int s = 0; s < 2; s++) {
      fill_spec(&fs.f[s], fl[s], 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk);
      vs.kind[s] = 0;
      vs.ptr[s] = fl[s];
    }
    if (open) fs.n = 0;
    run_fill(x, fs);
    run_vbc(x, vs);
  }
  stamp(x, 3);

  // ---- microphysics (cloudphy) on the future level ----
  if (cf.cphopt == 2) micro_stage(x, dtb, olf, open, 0);
  stamp(x, 4);

  // ---- Asselin filter (timeflt), then satadjst on present and future ----
  {
    float *xc[12] = {x->u[C], x->v[C], x->w[C], x->pp[C], x->ptp[C], x->q[C], x->q[C] + n, x->q[C] + 2 * n,
                     x->q[C] + 3 * n, x->q[C] + 4 * n, x->q[C] + 5 * n, x->nci[C]};
    float *xp[12] = {x->u[P], x->v[P], x->w[P], x->pp[P], x->ptp[P], x->q[P], x->q[P] + n, x->q[P] + 2 * n,
                     x->q[P] + 3 * n, x->q[P] + 4 * n, x->q[P] + 5 * n, x->nci[P]};
    float *xf[12] = {x->u[F], x->v[F], x->w[F], x->pp[F], x->ptp[F], x->q[F], x->q[F] + n, x->q[F] + 2 * n,
                     x->q[F] + 3 * n, x->q[F] + 4 * n, x->q[F] + 5 * n, x->nci[F]};
    if (!first && cf.cphopt == 2) {
      // pointwise filter and both adjustments fused per cell
      AsselinSat as;
      for (int f = 0; f < 12; f++) {
        as.x[f] = xc[f];
        as.xp[f] = xp[f];
        as.xf[f] = xf[f];
      }
      as.pbr = B[CRESS_PBR];
      as.ptbr = B[CRESS_PTBR];
      LAUNCH(x, KT_TIMEFLT, k_asselin_sat, flat(tiles(ni + 1, nj, 64, 4, nk)), b2, 0, tiles(ni + 1, nj, 64, 4, nk), g, c.m1fc2, c.filcoe, c.thresq, as);
      stamp(x, 5);
    } else {
      if (!first) {
        AsselinSet as;
        as.n = withq ? 12 : 6;
        const int kinds[12] = {1, 2, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int f = 0; f < 12; f++) {
          as.kind[f] = kinds[f];
          as.x[f] = xc[f];
          as.xp[f] = xp[f];
          as.xf[f] = xf[f];
        }
        LAUNCH(x, KT_TIMEFLT, k_timeflt, flat(tiles(ni + 1, nj, 64, 4, 12 * nk)), b2, 0, tiles(ni + 1, nj, 64, 4, 12 * nk), g, c, as);
      }
      stamp(x, 5);
      if (cf.cphopt == 2) {
        for (int lv = 0; lv < 2; lv++) {
          const int S = lv == 0 ? C : F;
          LAUNCH(x, KT_SATADJ, k_satadjst, flat(tiles(ni, nj - 1, 64, 4, nk - 1)), b2, 0, tiles(ni, nj - 1, 64, 4, nk - 1), g, c.thresq,
                 B[CRESS_PBR], B[CRESS_PTBR], x->pp[S], x->ptp[S], x->q[S], x->nci[S], n);
        }
      }
    }
  }
  stamp(x, 6);

  // ---- second halo rows for 4th-order advection (shift2nd) ----
  // s_shift2nd does nothing unless a 4th-order stencil reads the rows
  // (shift2nd.f90:266-268: advopt>=2 or the 4th-order smoothing; tubopt
  // and the exchange option 6 are not supported here).
  const bool need2 = x->cfg.advopt >= 2 || x->cfg.smtopt % 10 == 2 || x->cfg.smtopt % 10 == 3;
  for (int lv = 0; lv < 2 && open && need2; lv++) {
    // open sides: only the corners of the outer ring (shift2nd.f90:4590-4700)
    const int S = lv == 0 ? C : F;
    CornerSet cs;
    cs.n = 0;
    corner_spec(&cs, x->u[S], 0, ni + 1, 0, nj);
    corner_spec(&cs, x->v[S], 0, ni, 0, nj + 1);
    float *sc[10] = {x->w[S], x->pp[S], x->ptp[S], x->q[S], x->q[S] + n, x->q[S] + 2 * n,
                     x->q[S] + 3 * n, x->q[S] + 4 * n, x->q[S] + 5 * n, x->nci[S]};
    for (int f = 0; f < (withq ? 10 : 4); f++) corner_spec(&cs, sc[f], 0, ni, 0, nj);
    run_corners(x, cs);
  }
  for (int lv = 0; lv < 2 && !open && need2; lv++) {
    const int S = lv == 0 ? C : F;
    FillSet fs;
    fs.n = withq ? 12 : 6;
    fill_spec(&fs.f[0], x->u[S], 4, 0, ni - 3, ni + 1, 3, 0, nj - 3, nj, nk);
    fill_spec(&fs.f[1], x->v[S], 3, 0, ni - 3, ni, 4, 0, nj - 3, nj + 1, nk);
    float *sc[10] = {x->w[S], x->pp[S], x->ptp[S], x->q[S], x->q[S] + n, x->q[S] + 2 * n,
                     x->q[S] + 3 * n, x->q[S] + 4 * n, x->q[S] + 5 * n, x->nci[S]};
    for (int f = 0; f < 10; f++) fill_spec(&fs.f[2 + f], sc[f], 3, 0, ni - 3, ni, 3, 0, nj - 3, nj, nk);
    run_fill(x, fs);
  }
  stamp(x, 7);
  // ---- phasev: this step's dtb, dts set the next step's phase speeds ----
  x->obc_dtb = dtb;
  x->obc_dts = dts;
  x->obc_have = 1;
  // ---- swp2nxt: past <- present <- future by renaming ----
  x->rot = (x->rot + 1) % 3;
  stamp(x, 8);
  HIPCHK(hipGetLastError());
  return CRESS_OK;
}

int cress_large_step(cress_ctx *x, int nsteps, int first) {
  if (!x || nsteps < 0) return CRESS_EINVAL;
  for (int s = 0; s < nsteps; s++) {
    int rc = large_step(x, first && s == 0);
    if (rc) return rc;
    if (x->timing && x->tkern < 0) {
      HIPCHK(hipEventSynchronize(x->ev[NSTAGE]));
      flush_kernel_times(x);
      for (int t = 0; t < NSTAGE; t++) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, x->ev[t], x->ev[t + 1]);
        x->stage_ms[t] += ms;
      }
    } else if (x->timing && x->nkev > MAXEV / 2) {
      flush_kernel_times(x);
    }
  }
  // one family timed: its events are read once, after the last step
  if (x->timing && x->tkern >= 0) flush_kernel_times(x);
  return CRESS_OK;
}

// ---------------------------------------------------------------------------
// Ensemble: several contexts of any sizes stepped together.  Their large
// steps are captured once into one hipGraph whose branches are the members'
// own streams (fork and join by events on a master stream), so one graph
// launch advances every member by a cycle of steps and the members' kernels
// run side by side.  A captured step bakes in the buffers of the level
// rotation (and of the uf/vf alternation when nsstp is odd), so a cycle is
// 3 steps, or 6 when a member has odd nsstp: the rotation state after a
// cycle is the one it started from and the graph can be replayed as is.
// Steps that do not fill a cycle, and the cold-start first step, run
// eagerly; a graph is recaptured when the members' rotation state is not
// the one it was captured in.
// ---------------------------------------------------------------------------
// The host-side state of a member that its step launches depend on.
struct EnsState {
  int rot;
  float *u[3], *v[3], *ualt, *valt;
  float obc_dtb, obc_dts;
  int obc_have;
  int tune_ver;
  bool operator==(const EnsState &o) const {
    if (rot != o.rot || ualt != o.ualt || valt != o.valt || obc_have != o.obc_have) return false;
    if (tune_ver != o.tune_ver) return false;
    if (obc_have && (obc_dtb != o.obc_dtb || obc_dts != o.obc_dts)) return false;
    for (int l = 0; l < 3; l++)
      if (u[l] != o.u[l] || v[l] != o.v[l]) return false;
    return true;
  }
};

static EnsState ens_get(const cress_ctx *x) {
  EnsState t;
  t.rot = x->rot;
  for (int l = 0; l < 3; l++) {
    t.u[l] = x->u[l];
    t.v[l] = x->v[l];
  }
  t.ualt = x->ualt;
  t.valt = x->valt;
  t.obc_dtb = x->obc_dtb;
  t.obc_dts = x->obc_dts;
  t.obc_have = x->obc_have;
  t.tune_ver = x->tune_ver;
  return t;
}

static void ens_set(cress_ctx *x, const EnsState &t) {
  x->rot = t.rot;
  for (int l = 0; l < 3; l++) {
    x->u[l] = t.u[l];
    x->v[l] = t.v[l];
  }
  x->ualt = t.ualt;
  x->valt = t.valt;
  x->obc_dtb = t.obc_dtb;
  x->obc_dts = t.obc_dts;
  x->obc_have = t.obc_have;
}

struct cress_ens {
  std::vector<cress_ctx *> m;
  hipStream_t master;
  hipEvent_t fork;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec;
  bool have;
  int cycle;
  std::vector<EnsState> in, out;   // the members' state the graph starts from and leaves
};

int cress_ens_create(cress_ctx *const *members, int n, cress_ens **out) {
  if (!members || n < 1 || !out) return CRESS_EINVAL;
  cress_ens *e = new cress_ens();
  e->cycle = 3;
  e->have = false;
  for (int i = 0; i < n; i++) {
    bool dup = !members[i];
    for (int k = 0; k < i && !dup; k++) dup = members[k] == members[i];
    if (dup) {
      delete e;
      return CRESS_EINVAL;
    }
    e->m.push_back(members[i]);
    if (members[i]->cfg.nsstp % 2) e->cycle = 6;
  }
  if (hipStreamCreate(&e->master) != hipSuccess) {
    delete e;
    return CRESS_EHIP;
  }
  e->join.assign(n, nullptr);
  e->fork = nullptr;
  bool ok = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < n && ok; i++) ok = hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    for (hipEvent_t ev : e->join)
      if (ev) hipEventDestroy(ev);
    if (e->fork) hipEventDestroy(e->fork);
    hipStreamDestroy(e->master);
    delete e;
    return CRESS_EHIP;
  }
  *out = e;
  return CRESS_OK;
}

int cress_ens_destroy(cress_ens *e) {
  if (!e) return CRESS_EINVAL;
  hipStreamSynchronize(e->master);
  if (e->have) hipGraphExecDestroy(e->exec);
  for (hipEvent_t ev : e->join) hipEventDestroy(ev);
  hipEventDestroy(e->fork);
  hipStreamDestroy(e->master);
  delete e;
  return CRESS_OK;
}

// One cycle of steps of every member into a graph; the members' host-side
// state is put back to where it was, and the state the cycle leaves is kept
// for after each launch of the graph.
static int ens_capture(cress_ens *e) {
  if (e->have) {
    hipGraphExecDestroy(e->exec);
    e->have = false;
  }
  const size_t n = e->m.size();
  e->in.resize(n);
  e->out.resize(n);
  for (size_t i = 0; i < n; i++) {
    if (e->m[i]->timing) return CRESS_EINVAL;   // events are not captured
    e->in[i] = ens_get(e->m[i]);
  }
  HIPCHK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
  int rc = hipEventRecord(e->fork, e->master) == hipSuccess ? CRESS_OK : CRESS_EHIP;
  for (size_t i = 0; i < n && rc == CRESS_OK; i++) {
    cress_ctx *x = e->m[i];
    if (hipStreamWaitEvent(x->stream, e->fork, 0) != hipSuccess) rc = CRESS_EHIP;
    for (int s = 0; s < e->cycle && rc == CRESS_OK; s++) rc = large_step(x, 0);
    if (rc == CRESS_OK && hipEventRecord(e->join[i], x->stream) != hipSuccess) rc = CRESS_EHIP;
    if (rc == CRESS_OK && hipStreamWaitEvent(e->master, e->join[i], 0) != hipSuccess) rc = CRESS_EHIP;
  }
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(e->master, &graph);
  for (size_t i = 0; i < n; i++) {
    e->out[i] = ens_get(e->m[i]);
    ens_set(e->m[i], e->in[i]);
  }
  if (rc != CRESS_OK || end != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return rc != CRESS_OK ? rc : CRESS_EHIP;
  }
  const hipError_t ie = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ie != hipSuccess) return CRESS_EHIP;
  e->have = true;
  return CRESS_OK;
}

static bool ens_ready(const cress_ens *e) {
  if (!e->have) return false;
  for (size_t i = 0; i < e->m.size(); i++)
    if (!(ens_get(e->m[i]) == e->in[i])) return false;
  return true;
}

int cress_ens_step(cress_ens *e, int nsteps, int first) {
  if (!e || nsteps < 0) return CRESS_EINVAL;
  int done = 0;
  if (first && nsteps > 0) {
    for (cress_ctx *x : e->m) {
      const int rc = large_step(x, 1);
      if (rc) return rc;
    }
    done = 1;
  }
  while (nsteps - done >= e->cycle) {
    // every member's earlier work is ordered before the graph
    for (size_t i = 0; i < e->m.size(); i++) {
      HIPCHK(hipEventRecord(e->join[i], e->m[i]->stream));
      HIPCHK(hipStreamWaitEvent(e->master, e->join[i], 0));
    }
    if (!ens_ready(e)) {
      const int rc = ens_capture(e);
      if (rc) return rc;
    }
    HIPCHK(hipGraphLaunch(e->exec, e->master));
    for (size_t i = 0; i < e->m.size(); i++) ens_set(e->m[i], e->out[i]);
    // and the members' later work after it
    HIPCHK(hipEventRecord(e->fork, e->master));
    for (cress_ctx *x : e->m) HIPCHK(hipStreamWaitEvent(x->stream, e->fork, 0));
    done += e->cycle;
  }
  for (; done < nsteps; done++)
    for (cress_ctx *x : e->m) {
      const int rc = large_step(x, 0);
      if (rc) return rc;
    }
  return CRESS_OK;
}

int cress_ens_sync(cress_ens *e) {
  if (!e) return CRESS_EINVAL;
  HIPCHK(hipStreamSynchronize(e->master));
  for (cress_ctx *x : e->m) HIPCHK(hipStreamSynchronize(x->stream));
  return CRESS_OK;
}

// ---------------------------------------------------------------------------
// Tier 2 shims
// ---------------------------------------------------------------------------
// Mass and energy budgets of one level in binary64 (diag.hip); out gets
// CRESS_NDIAG values in the order of include/cress_hip.h.
int cress_diag(cress_ctx *x, int level, double *out) {
  if (!x || !out || level < 0 || level > 2) return CRESS_EINVAL;
  const Grid g = x->g;
  const int nk = g.nk, s = (x->rot + level) % 3;
  const long n = x->n;
  double *part = x->diag, *res = x->diag + (size_t)kDiagBlocks * DIAG_N, *zs = res + DIAG_N;
  // heights of the scalar levels over flat terrain from the per-level
  // Jacobian (z at w-level 2 is the surface): z_w(k+1) = z_w(k) + dz jcb(k),
  // z_s(k) = z_w(k) + dz jcb(k)/2; NaN when jcb varies horizontally
  std::vector<double> z((size_t)nk + 1, 0.0);
  const bool jcbu = (g.umask >> CRESS_JCB) & 1u;
  double zw = 0.0;
  for (int k = 2; k <= nk - 1; k++) {
    const double jk = jcbu ? (double)x->tabh[(size_t)CRESS_JCB * nk + k - 1] : NAN;
    z[k] = zw + 0.5 * x->cfg.dz * jk;
    zw += x->cfg.dz * jk;
  }
  HIPCHK(hipMemcpyAsync(zs, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
  DiagArgs a;
  a.u = x->u[s];
  a.v = x->v[s];
  a.w = x->w[s];
  a.pp = x->pp[s];
  a.ptp = x->ptp[s];
  a.q = x->q[s];
  a.qn = n;
  a.rst = x->base[CRESS_RST];
  a.jcb = x->base[CRESS_JCB];
  a.pbr = x->base[CRESS_PBR];
  a.ptbr = x->base[CRESS_PTBR];
  a.prwtr = x->prwtr;
  a.price = x->price;
  a.zs = zs;
  a.dx = x->cfg.dx;
  a.dy = x->cfg.dy;
  a.dz = x->cfg.dz;
  hipLaunchKernelGGL(k_diag_cells, dim3(kDiagBlocks), dim3(256), 0, x->stream, g, a, part);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, x->stream, (const double *)part, kDiagBlocks, res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, res, DIAG_N * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return CRESS_OK;
}

// ---------------------------------------------------------------------------
// Launch-shape autotuner.  Each tunable kernel family is timed on this
// context's own grid for every candidate shape (rows per workgroup, tile
// order, halo layout, LDS occupancy cap), and the fastest is kept.  The
// candidates write only scratch arrays and the future level, which the next
// large step forms again from the past level, so tuning may run between any
// two steps; the results do not depend on the shapes (tests show them
// bitwise equal).  Results are cached per process by grid size and options.
namespace {
constexpr int kTuneInts = 9;

void tune_to_ints(const LaunchTune &t, int *v) {
  const int a[kTuneInts] = {t.ptuv_by, t.ptuv_slots, t.ptuv_order, t.col_by, t.col_lds,
                            t.fs_by,   t.fs_order,   t.ff_by,      t.ff_order};
  for (int i = 0; i < kTuneInts; i++) v[i] = a[i];
}

bool tune_valid(const LaunchTune &t) {
  auto in = [](int v, std::initializer_list<int> set) {
    for (int s : set)
      if (v == s) return true;
    return false;
  };
  return in(t.ptuv_by, {1, 2, 4, 8, 16}) && in(t.ptuv_slots, {0, 1}) && in(t.ptuv_order, {0, 1, 2}) &&
         in(t.col_by, {1, 2, 4}) && t.col_lds >= 0 && t.col_lds <= 65536 && in(t.fs_by, {1, 2, 4, 8}) &&
         in(t.fs_order, {0, 1, 2}) && in(t.ff_by, {1, 2, 4, 8, 16}) && in(t.ff_order, {0, 1, 2});
}

std::mutex tune_mu;
std::map<std::vector<long>, LaunchTune> tune_cache;

std::vector<long> tune_key(const cress_ctx *x) {
  const cress_config &c = x->cfg;
  return {c.ni, c.nj, c.nk, c.cphopt, c.advopt, c.smtopt, c.wbc, c.divopt, c.buyopt, (long)x->g.umask};
}
}  // namespace

int cress_set_launch(cress_ctx *x, const int *shape, int n) {
  if (!x || !shape || n != kTuneInts) return CRESS_EINVAL;
  LaunchTune t;
  t.ptuv_by = shape[0];
  t.ptuv_slots = shape[1];
  t.ptuv_order = shape[2];
  t.col_by = shape[3];
  t.col_lds = shape[4];
  t.fs_by = shape[5];
  t.fs_order = shape[6];
  t.ff_by = shape[7];
  t.ff_order = shape[8];
  if (!tune_valid(t)) return CRESS_EINVAL;
  x->tune = t;
  x->tune_ver++;
  return CRESS_OK;
}

int cress_get_launch(const cress_ctx *x, int *shape, int n) {
  if (!x || !shape || n < kTuneInts) return CRESS_EINVAL;
  tune_to_ints(x->tune, shape);
  return CRESS_OK;
}

int cress_autotune(cress_ctx *x, int *shape, int n) {
  if (!x || (shape && n < kTuneInts)) return CRESS_EINVAL;
  if (x->timing) return CRESS_EINVAL;
  const std::vector<long> key = tune_key(x);
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_cache.find(key);
    if (it != tune_cache.end()) {
      x->tune = it->second;
      x->tune_ver++;
      if (shape) tune_to_ints(x->tune, shape);
      return CRESS_OK;
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) hipEventDestroy(e0);
    return CRESS_EHIP;
  }
  int rc = CRESS_OK;
  // mean device time of one launch sequence: one untimed run, one timed
  // run to size the repetitions (about 2 ms in all, 3 to 40 runs), then
  // the timed repetitions
  auto time_ms = [&](auto launch) -> float {
    launch();
    float ms = 0.0f;
    if (hipEventRecord(e0, x->stream) != hipSuccess) rc = CRESS_EHIP;
    launch();
    if (hipEventRecord(e1, x->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = CRESS_EHIP;
    int reps = ms > 0.0f ? (int)(2.0f / ms) : 40;
    reps = reps < 3 ? 3 : (reps > 40 ? 40 : reps);
    if (hipEventRecord(e0, x->stream) != hipSuccess) rc = CRESS_EHIP;
    for (int r = 0; r < reps; r++) launch();
    if (hipEventRecord(e1, x->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = CRESS_EHIP;
    return ms / (float)reps;
  };
  LaunchTune best = x->tune;
  const float dts = (float)x->cfg.dtsml;
  // fused ptil/stepuv
  {
    const AcPtUv pu = ptuv_args(x, dts);
    float tmin = 1e30f;
    for (int by : {4, 8, 16})
      for (int slots : {0, 1})
        for (int order : {1, 2}) {
          LaunchTune t = best;
          t.ptuv_by = by;
          t.ptuv_slots = slots;
          t.ptuv_order = order;
          const float ms = time_ms([&] { launch_ptuv(x, t, pu); });
          if (ms < tmin) {
            tmin = ms;
            best.ptuv_by = by;
            best.ptuv_slots = slots;
            best.ptuv_order = order;
          }
        }
  }
  // column kernel
  {
    const AcCol ac = col_args(x, dts);
    float tmin = 1e30f;
    for (int by : {1, 2, 4})
      for (int lds : {0, 4096, 8192, 16384}) {
        LaunchTune t = best;
        t.col_by = by;
        t.col_lds = lds;
        const float ms = time_ms([&] { launch_col(x, t, ac); });
        if (ms < tmin) {
          tmin = ms;
          best.col_by = by;
          best.col_lds = lds;
        }
      }
  }
  // scalar forcing, then the u/v, w and p' forcing together
  {
    float tmin = 1e30f;
    for (int by : {2, 4, 8})
      for (int order : {1, 2}) {
        LaunchTune t = best;
        t.fs_by = by;
        t.fs_order = order;
        const float ms = time_ms([&] { forcing_stage(x, t, FORCE_S); });
        if (ms < tmin) {
          tmin = ms;
          best.fs_by = by;
          best.fs_order = order;
        }
      }
  }
  {
    float tmin = 1e30f;
    for (int by : {2, 4, 8})
      for (int order : {1, 2}) {
        LaunchTune t = best;
        t.ff_by = by;
        t.ff_order = order;
        const float ms = time_ms([&] { forcing_stage(x, t, FORCE_UV | FORCE_W | FORCE_P); });
        if (ms < tmin) {
          tmin = ms;
          best.ff_by = by;
          best.ff_order = order;
        }
      }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (hipGetLastError() != hipSuccess) rc = CRESS_EHIP;
  if (rc) return rc;
  x->tune = best;
  x->tune_ver++;
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache[key] = best;
  }
  if (shape) tune_to_ints(best, shape);
  return CRESS_OK;
}

// ---------------------------------------------------------------------------
// Base state from a 1-D sounding in binary64 on the device (setup.hip).
int cress_set_base(cress_ctx *x, const cress_sounding *s) {
  if (!x || !s || s->nlev < 1 || !s->z || !s->u || !s->v || !s->p || !s->pt || !s->qv) return CRESS_EINVAL;
  const cress_config &cf = x->cfg;
  if (cf.sthopt != 0 && cf.sthopt != 1) return CRESS_EUNSUPPORTED;
  if (cf.sthopt == 1 && !(s->dzmin > 0.0)) return CRESS_EINVAL;
  for (int k = 0; k < s->nlev; k++)
    if (!(s->z[k + 1] > s->z[k])) return CRESS_EINVAL;
  const Grid g = x->g;
  const int nk = g.nk, ld = nk + 1;
  const size_t nl = (size_t)s->nlev + 1;
  if (!x->setupcol) {
    float *t = nullptr;
    int rc = dalloc(x, &t, 2 * (size_t)SC_N * ld);
    if (rc) return rc;
    x->setupcol = (double *)t;
  }
  double *snd = nullptr;
  if (hipMalloc(&snd, 6 * nl * sizeof(double)) != hipSuccess) return CRESS_ENOMEM;
  const double *src[6] = {s->z, s->u, s->v, s->p, s->pt, s->qv};
  int rc = CRESS_OK;
  for (int c = 0; c < 6 && !rc; c++)
    if (hipMemcpyAsync(snd + c * nl, src[c], nl * sizeof(double), hipMemcpyHostToDevice, x->stream) != hipSuccess)
      rc = CRESS_EHIP;
  SetupArgs a;
  a.nk = nk;
  a.sthopt = cf.sthopt;
  a.bbc = cf.bbc;
  a.tbc = cf.tbc;
  a.dz = cf.dz;
  a.zsfc = s->zsfc;
  a.dzmin = s->dzmin;
  a.layer1 = s->layer1;
  a.layer2 = s->layer2;
  a.botbar = s->botbar;
  a.vspbar = s->vspbar;
  a.nlev = s->nlev;
  a.z = snd;
  a.u = snd + nl;
  a.v = snd + 2 * nl;
  a.p = snd + 3 * nl;
  a.pt = snd + 4 * nl;
  a.qv = snd + 5 * nl;
  SetupFill f;
  f.col = x->setupcol;
  for (int id = 0; id < CRESS_NBASE; id++) f.base[id] = x->base[id];
  f.tab = x->tab;
  f.rbct = x->rbct;
  f.fc = x->fc;
  f.fcor = s->fc;
  if (!rc) {
    hipLaunchKernelGGL(k_setup_col, dim3(1), dim3(64), 0, x->stream, a, x->setupcol);
    const long cells = g.nxy * (long)nk;
    hipLaunchKernelGGL(k_setup_fill, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, x->stream, g, f);
    if (hipGetLastError() != hipSuccess) rc = CRESS_EHIP;
  }
  std::vector<double> zph(ld);
  if (!rc && (hipMemcpyAsync(zph.data(), x->setupcol + SC_ZPH * ld, ld * sizeof(double), hipMemcpyDeviceToHost,
                             x->stream) != hipSuccess ||
              hipMemcpyAsync(x->tabh.data(), x->tab, x->tabh.size() * sizeof(float), hipMemcpyDeviceToHost,
                             x->stream) != hipSuccess ||
              hipStreamSynchronize(x->stream) != hipSuccess))
    rc = CRESS_EHIP;
  hipFree(snd);
  if (rc) return rc;
  // every base field is now one value per level; vspdmp's lowest sponge
  // level: one below the first zeta level (from k=3) above botbar
  x->g.umask = (1u << CRESS_NBASE) - 1u;
  int ksp0 = 2;
  for (int k = 3; k <= nk; k++)
    if (zph[k] > s->botbar) {
      ksp0 = k - 1;
      break;
    }
  x->cfg.ksp0[0] = 2;
  x->cfg.ksp0[1] = ksp0;
  x->o.ksp0 = ksp0;
  return refresh_accoef(x);
}

int cress_base_columns(const cress_ctx *x, double *out, int ncol) {
  if (!x || !out || ncol < 1 || ncol > SC_N) return CRESS_EINVAL;
  if (!x->setupcol) return CRESS_ERANGE;
  const size_t ld = (size_t)x->g.nk + 1;
  if (hipMemcpy(out, x->setupcol, (size_t)ncol * ld * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return CRESS_EHIP;
  return CRESS_OK;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

static int to_dev(DevBuf &b, const float *h, size_t n) {
  if (hipMalloc(&b.p, n * sizeof(float)) != hipSuccess) return CRESS_ENOMEM;
  if (h && hipMemcpy(b.p, h, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return CRESS_EHIP;
  return CRESS_OK;
}

int cress_gaussel(int istr, int iend, int jstr, int jend, int kstr, int kend, int ni, int nj, int nk,
                  float *rr, float *ss, float *tt, float *ff) {
  if (!rr || !ss || !tt || !ff || ni < 1 || nj < 1 || nk < 1) return CRESS_EINVAL;
  if (istr < 0 || iend > ni + 1 || jstr < 0 || jend > nj + 1 || kstr < 1 || kend > nk || kstr > kend)
    return CRESS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(ni, nj, nk);
  const size_t n = (size_t)g.n();
  DevBuf a, b, c, d;
  int rc;
  if ((rc = to_dev(a, rr, n)) || (rc = to_dev(b, ss, n)) || (rc = to_dev(c, tt, n)) || (rc = to_dev(d, ff, n)))
    return rc;
  const int ncol = (iend - istr + 1) * (jend - jstr + 1);
  hipLaunchKernelGGL(k_gaussel, dim3((ncol + 63) / 64), dim3(64), 0, 0, g, istr, iend, jstr, jend, kstr, kend,
                     (const float *)a.p, (float *)b.p, (const float *)c.p, (float *)d.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(ss, b.p, n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ff, d.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// ---------------------------------------------------------------------------
// cress_gaussel without allocation: the four column arrays are staged in a
// file-scope device array instead of hipMalloc'ed buffers.  Each GPU holds
// its own copy of `scratch` (CRESS_GAUSSEL_SCRATCH_FLOATS floats, 32 MiB)
// from the library's first use on it until the process ends.  In host code
// the array's name is not a device address; its address on each device is
// looked up once (hipGetSymbolAddress, with that device current) and kept in
// scratch_dev[].  Calls overlapping on one GPU share the array and race.
static __device__ __align__(256) float scratch[CRESS_GAUSSEL_SCRATCH_FLOATS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float *scratch_dev[kMaxDevices];

int scratch_ptr(float **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CRESS_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return CRESS_ENODEV;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CRESS_EHIP;
    scratch_dev[dev] = (float *)p;
  }
  *out = scratch_dev[dev];
  return CRESS_OK;
}
}  // namespace

int cress_gaussel_noalloc(int istr, int iend, int jstr, int jend, int kstr, int kend, int ni, int nj, int nk,
                          float *rr, float *ss, float *tt, float *ff) {
  if (!rr || !ss || !tt || !ff || ni < 1 || nj < 1 || nk < 1) return CRESS_EINVAL;
  if (istr < 0 || iend > ni + 1 || jstr < 0 || jend > nj + 1 || kstr < 1 || kend > nk || kstr > kend)
    return CRESS_EINVAL;
  const Grid g = make_grid(ni, nj, nk);
  const size_t n = (size_t)g.n();
  if (n > (size_t)CRESS_GAUSSEL_SCRATCH_FLOATS / 4) return CRESS_ETOOBIG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  float *base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  // the four arrays back to back; every element is overwritten by its
  // upload, so nothing of the previous call's contents is read
  float *a = base, *b = base + n, *c = base + 2 * n, *d = base + 3 * n;
  HIPCHK(hipMemcpy(a, rr, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b, ss, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c, tt, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d, ff, n * sizeof(float), hipMemcpyHostToDevice));
  const int ncol = (iend - istr + 1) * (jend - jstr + 1);
  hipLaunchKernelGGL(k_gaussel, dim3((ncol + 63) / 64), dim3(64), 0, 0, g, istr, iend, jstr, jend, kstr, kend,
                     (const float *)a, b, (const float *)c, d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(ss, b, n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ff, d, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

int cress_advs(const cress_config *cfg, const float *rstxu, const float *rstxv, const float *rstxwc,
               const float *s, float *sfrc) {
  if (!cfg || !rstxu || !rstxv || !rstxwc || !s || !sfrc) return CRESS_EINVAL;
  if (cfg->ni < 7 || cfg->nj < 7 || cfg->nk < 7 || cfg->advopt < 1 || cfg->advopt > 3) return CRESS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  Consts c;
  fold_consts(cfg, &c);
  Opts o = {};
  o.advopt = cfg->advopt;
  o.iw = cfg->wbc == -1 ? 0 : 1;
  o.ie = cfg->ebc == -1 ? 0 : 1;
  o.js = cfg->sbc == -1 ? 0 : 1;
  o.jn = cfg->nbc == -1 ? 0 : 1;
  DevBuf a, b, d, e, f;
  int rc;
  if ((rc = to_dev(a, rstxu, n)) || (rc = to_dev(b, rstxv, n)) || (rc = to_dev(d, rstxwc, n)) ||
      (rc = to_dev(e, s, n)) || (rc = to_dev(f, sfrc, n)))
    return rc;
  ScalarIn si = {};
  si.s[0] = (const float *)e.p;
  si.frc[0] = (float *)f.p;
  si.rstxu = (const float *)a.p;
  si.rstxv = (const float *)b.p;
  si.rstxwc = (const float *)d.p;
  hipLaunchKernelGGL(k_advs_only, grid2(g.ni - 3, g.nj - 3, 64, 4, g.nk - 3), dim3(64, 4), 0, 0, g, o, c, si);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(sfrc, f.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

int cress_advuvw(const cress_config *cfg, const float *rstxu, const float *rstxv, const float *rstxwc,
                 const float *u, const float *v, const float *w, float *ufrc, float *vfrc, float *wfrc) {
  if (!cfg || !rstxu || !rstxv || !rstxwc || !u || !v || !w || !ufrc || !vfrc || !wfrc) return CRESS_EINVAL;
  if (cfg->ni < 7 || cfg->nj < 7 || cfg->nk < 7 || cfg->advopt < 1 || cfg->advopt > 3) return CRESS_EINVAL;
  const int sides[4] = {cfg->wbc, cfg->ebc, cfg->sbc, cfg->nbc};
  for (int s = 0; s < 4; s++)
    if (sides[s] != -1 && sides[s] != 7) return CRESS_EUNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  Consts c;
  fold_consts(cfg, &c);
  Opts o = {};
  o.advopt = cfg->advopt;
  // the 4th-order terms stop one point short of a non-periodic side
  // (sethalo.f90:536-580)
  o.iw = cfg->wbc == -1 ? 0 : 1;
  o.ie = cfg->ebc == -1 ? 0 : 1;
  o.js = cfg->sbc == -1 ? 0 : 1;
  o.jn = cfg->nbc == -1 ? 0 : 1;
  DevBuf a, b, d, e, f, h, pu, pv, pw;
  int rc;
  if ((rc = to_dev(a, rstxu, n)) || (rc = to_dev(b, rstxv, n)) || (rc = to_dev(d, rstxwc, n)) ||
      (rc = to_dev(e, u, n)) || (rc = to_dev(f, v, n)) || (rc = to_dev(h, w, n)) || (rc = to_dev(pu, ufrc, n)) ||
      (rc = to_dev(pv, vfrc, n)) || (rc = to_dev(pw, wfrc, n)))
    return rc;
  hipLaunchKernelGGL(k_advuvw_only, grid2(g.ni - 2, g.nj - 2, 64, 4, g.nk - 2), dim3(64, 4), 0, 0, g, o, c,
                     (const float *)e.p, (const float *)f.p, (const float *)h.p, (const float *)a.p,
                     (const float *)b.p, (const float *)d.p, (float *)pu.p, (float *)pv.p, (float *)pw.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(ufrc, pu.p, n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vfrc, pv.p, n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(wfrc, pw.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// s_hevi for one substep: a context of the caller's configuration holds
// the arrays, the substep of the large step runs once, and the results come
// back (periodic sides).
int cress_hevi(const cress_config *cfg, float dts, const float *jcb, const float *jcb8u, const float *jcb8v,
               const float *jcb8w, const float *rbr, const float *rst, const float *rst8u, const float *rst8v,
               const float *rst8w, const float *rcsq, const float *ufrc, const float *vfrc, const float *wfrc,
               const float *pfrc, float *uf, float *vf, float *wf, float *wc, float *ppf) {
  const float *in[14] = {jcb, jcb8u, jcb8v, jcb8w, rbr, rst, rst8u, rst8v, rst8w, rcsq, ufrc, vfrc, wfrc, pfrc};
  for (const float *a : in)
    if (!a) return CRESS_EINVAL;
  if (!cfg || !uf || !vf || !wf || !wc || !ppf || !(dts > 0.0f)) return CRESS_EINVAL;
  if (!(cfg->wbc == -1 && cfg->ebc == -1 && cfg->sbc == -1 && cfg->nbc == -1)) return CRESS_EUNSUPPORTED;
  cress_ctx *x = nullptr;
  int rc = cress_ctx_create(cfg, &x);
  if (rc) return rc;
  const int ids[10] = {CRESS_JCB, CRESS_JCB8U, CRESS_JCB8V, CRESS_JCB8W, CRESS_RBR,
                       CRESS_RST, CRESS_RST8U, CRESS_RST8V, CRESS_RST8W, CRESS_RCSQ};
  for (int f = 0; f < 10 && !rc; f++) rc = cress_upload(x, ids[f], 0, in[f]);
  if (!rc) rc = cress_upload(x, CRESS_U, 2, uf);
  if (!rc) rc = cress_upload(x, CRESS_V, 2, vf);
  if (!rc) rc = cress_upload(x, CRESS_W, 2, wf);
  if (!rc) rc = cress_upload(x, CRESS_PP, 2, ppf);
  float *scr[7] = {x->ufrc, x->vfrc, x->wfrc, x->pfrc, x->wc, x->ualt, x->valt};
  const float *src[7] = {ufrc, vfrc, wfrc, pfrc, wc, uf, vf};
  for (int f = 0; f < 7 && !rc; f++) rc = copy_field(x, scr[f], src[f], true);
  if (!rc) {
    ac_substep(x, dts, nullptr);
    if (hipGetLastError() != hipSuccess) rc = CRESS_EHIP;
  }
  if (!rc) rc = cress_download(x, CRESS_U, 2, uf);
  if (!rc) rc = cress_download(x, CRESS_V, 2, vf);
  if (!rc) rc = cress_download(x, CRESS_W, 2, wf);
  if (!rc) rc = cress_download(x, CRESS_PP, 2, ppf);
  if (!rc) rc = copy_field(x, x->wc, wc, false);
  if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) rc = CRESS_EHIP;
  cress_ctx_destroy(x);
  return rc;
}

// s_steps on host arrays: a context of the caller's configuration holds
// them; the large step's own scalar leapfrog (k_steps), fills and vbcs run
// once (periodic sides).
int cress_steps(const cress_config *cfg, float dtb, const float *rst, const float *ptpp, const float *ptfrc,
                const float *qvp, const float *qvfrc, const float *qwtrp, const float *qicep, const float *nicep,
                float *qwtrf, float *qicef, float *nicef, float *ptpf, float *qvf, float *dtdrst) {
  if (!cfg || !rst || !ptpp || !ptfrc || !qvp || !qvfrc || !ptpf || !qvf || !dtdrst || !(dtb > 0.0f))
    return CRESS_EINVAL;
  const bool withq = cfg->cphopt == 2;
  if (withq && (!qwtrp || !qicep || !nicep || !qwtrf || !qicef || !nicef)) return CRESS_EINVAL;
  if (!(cfg->wbc == -1 && cfg->ebc == -1 && cfg->sbc == -1 && cfg->nbc == -1)) return CRESS_EUNSUPPORTED;
  cress_ctx *x = nullptr;
  int rc = cress_ctx_create(cfg, &x);
  if (rc) return rc;
  const Grid g = x->g;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const size_t h = (size_t)(ni + 2) * (nj + 2) * nk;   // one host field
  const int P = (x->rot + 0) % 3, F = (x->rot + 2) % 3;
  rc = cress_upload(x, CRESS_RST, 0, rst);
  if (!rc) rc = cress_upload(x, CRESS_PTP, 0, ptpp);
  if (!rc) rc = cress_upload(x, CRESS_QV, 0, qvp);
  // the cells steps does not write keep what the caller's arrays hold
  if (!rc) rc = cress_upload(x, CRESS_PTP, 2, ptpf);
  if (!rc) rc = cress_upload(x, CRESS_QV, 2, qvf);
  if (!rc) rc = copy_field(x, x->ptfrc, ptfrc, true);
  if (!rc) rc = copy_field(x, x->qvfrc, qvfrc, true);
  const int sid[5] = {CRESS_QC, CRESS_QR, CRESS_QI, CRESS_QS, CRESS_QG};
  const float *sp[5] = {qwtrp, qwtrp + h, qicep, qicep + h, qicep + 2 * h};
  float *sf[5] = {qwtrf, qwtrf + h, qicef, qicef + h, qicef + 2 * h};
  for (int s = 0; s < 5 && withq && !rc; s++) {
    rc = cress_upload(x, sid[s], 0, sp[s]);
    if (!rc) rc = cress_upload(x, sid[s], 2, sf[s]);
  }
  if (withq && !rc) rc = cress_upload(x, CRESS_NCI, 0, nicep);
  if (withq && !rc) rc = cress_upload(x, CRESS_NCI, 2, nicef);
  float *ddev = nullptr;
  if (!rc && hipMalloc(&ddev, (size_t)g.n() * sizeof(float)) != hipSuccess) rc = CRESS_ENOMEM;
  if (!rc) rc = copy_field(x, ddev, dtdrst, true);
  if (!rc) {
    const long n = x->n;
    StepsIn sti;
    sti.ptpp = x->ptp[P];
    sti.ptfrc = x->ptfrc;
    sti.qp = x->q[P];
    sti.qvfrc = x->qvfrc;
    sti.ncip = x->nci[P];
    sti.ptpf = x->ptp[F];
    sti.qf = x->q[F];
    sti.ncif = x->nci[F];
    sti.qn = n;
    const Tiles ts = tiles(ni - 1, nj - 3, 64, 4, nk - 3);
    const float dtb2 = 2.0f * dtb;
    float **B = x->base;
    hipLaunchKernelGGL(k_steps, flat(ts), dim3(64, 4), 0, x->stream, ts, g, dtb2, B[CRESS_RST], sti, withq ? 1 : 0);
    hipLaunchKernelGGL(k_dtdrst, grid2(ni - 3, nj - 3, 64, 4, nk - 3), dim3(64, 4), 0, x->stream, g, dtb2,
                       B[CRESS_RST], ddev);
    FillSet fs;
    VbcSet vs;
    float *fl[8] = {x->ptp[F], x->q[F], x->q[F] + n, x->q[F] + 2 * n, x->q[F] + 3 * n, x->q[F] + 4 * n,
                    x->q[F] + 5 * n, x->nci[F]};
    fs.n = withq ? 8 : 2;
    vs.n = fs.n;
    for (int s = 0; s < fs.n; s++) {
      fill_spec(&fs.f[s], fl[s], 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk);
      vs.kind[s] = 0;
      vs.ptr[s] = fl[s];
    }
    run_fill(x, fs);
    run_vbc(x, vs);
    if (hipGetLastError() != hipSuccess) rc = CRESS_EHIP;
  }
  if (!rc) rc = cress_download(x, CRESS_PTP, 2, ptpf);
  if (!rc) rc = cress_download(x, CRESS_QV, 2, qvf);
  for (int s = 0; s < 5 && withq && !rc; s++) rc = cress_download(x, sid[s], 2, sf[s]);
  if (withq && !rc) rc = cress_download(x, CRESS_NCI, 2, nicef);
  if (!rc) rc = copy_field(x, ddev, dtdrst, false);
  if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) rc = CRESS_EHIP;
  if (ddev) hipFree(ddev);
  cress_ctx_destroy(x);
  return rc;
}

// The TKE lines of s_steps: a dry context of the caller's configuration
// holds rst, and its theta' slots hold tkep (past) and tkef (future); the
// future level then takes the same halo fills and vbcs as theta'.
int cress_stepstke(const cress_config *cfg, float dtb, const float *rst, const float *tkep, float *tkef) {
  if (!cfg || !rst || !tkep || !tkef || !(dtb > 0.0f)) return CRESS_EINVAL;
  if (!(cfg->wbc == -1 && cfg->ebc == -1 && cfg->sbc == -1 && cfg->nbc == -1)) return CRESS_EUNSUPPORTED;
  cress_config dry = *cfg;
  dry.cphopt = 0;
  cress_ctx *x = nullptr;
  int rc = cress_ctx_create(&dry, &x);
  if (rc) return rc;
  const Grid g = x->g;
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const int P = (x->rot + 0) % 3, F = (x->rot + 2) % 3;
  rc = cress_upload(x, CRESS_RST, 0, rst);
  if (!rc) rc = cress_upload(x, CRESS_PTP, 0, tkep);
  if (!rc) rc = cress_upload(x, CRESS_PTP, 2, tkef);
  if (!rc) {
    hipLaunchKernelGGL(k_steps_tke, grid2(ni - 3, nj - 3, 64, 4, nk - 3), dim3(64, 4), 0, x->stream, g, 2.0f * dtb,
                       x->base[CRESS_RST], x->ptp[P], x->ptp[F]);
    FillSet fs;
    VbcSet vs;
    fs.n = vs.n = 1;
    fill_spec(&fs.f[0], x->ptp[F], 2, 1, ni - 2, ni - 1, 2, 1, nj - 2, nj - 1, nk);
    vs.kind[0] = 0;
    vs.ptr[0] = x->ptp[F];
    run_fill(x, fs);
    run_vbc(x, vs);
    if (hipGetLastError() != hipSuccess) rc = CRESS_EHIP;
  }
  if (!rc) rc = cress_download(x, CRESS_PTP, 2, tkef);
  if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) rc = CRESS_EHIP;
  cress_ctx_destroy(x);
  return rc;
}

// s_cloudphy on host arrays for cphopt=2: a context of the caller's
// configuration holds them and the large step's microphysics stage runs
// once on the given future level.
int cress_cloudphy(const cress_config *cfg, float dtb, const float *jcb, const float *pbr, const float *ptbr,
                   const float *rbr, const float *rst, const float *ppp, const float *ptpp, const float *qvp,
                   const float *qwtrp, const float *qicep, const float *nicep, float *ptpf, float *qvf,
                   float *qwtrf, float *qicef, float *nicef, float *prwtr, float *price) {
  const float *in[12] = {jcb, pbr, ptbr, rbr, rst, ppp, ptpp, qvp, qwtrp, qicep, nicep, ptpf};
  for (const float *a : in)
    if (!a) return CRESS_EINVAL;
  if (!cfg || !qvf || !qwtrf || !qicef || !nicef || !prwtr || !price || !(dtb > 0.0f)) return CRESS_EINVAL;
  if (cfg->cphopt != 2) return CRESS_EUNSUPPORTED;
  cress_ctx *x = nullptr;
  int rc = cress_ctx_create(cfg, &x);
  if (rc) return rc;
  const size_t h = (size_t)(cfg->ni + 2) * (cfg->nj + 2) * cfg->nk;
  const int bid[5] = {CRESS_JCB, CRESS_PBR, CRESS_PTBR, CRESS_RBR, CRESS_RST};
  for (int f = 0; f < 5 && !rc; f++) rc = cress_upload(x, bid[f], 0, in[f]);
  const int fid[8] = {CRESS_PTP, CRESS_QV, CRESS_QC, CRESS_QR, CRESS_QI, CRESS_QS, CRESS_QG, CRESS_NCI};
  const float *past[8] = {ptpp, qvp, qwtrp, qwtrp + h, qicep, qicep + h, qicep + 2 * h, nicep};
  float *fut[8] = {ptpf, qvf, qwtrf, qwtrf + h, qicef, qicef + h, qicef + 2 * h, nicef};
  if (!rc) rc = cress_upload(x, CRESS_PP, 0, ppp);
  for (int f = 0; f < 8 && !rc; f++) {
    rc = cress_upload(x, fid[f], 0, past[f]);
    if (!rc) rc = cress_upload(x, fid[f], 2, fut[f]);
  }
  if (!rc) rc = cress_upload(x, CRESS_PRWTR, 0, prwtr);
  if (!rc) rc = cress_upload(x, CRESS_PRICE, 0, price);
  if (!rc) {
    micro_stage(x, dtb, ObcLf{0.0f, 0.0f, 0.0f, 0.0f}, 0, 1);
    if (hipGetLastError() != hipSuccess) rc = CRESS_EHIP;
  }
  for (int f = 0; f < 8 && !rc; f++) rc = cress_download(x, fid[f], 2, fut[f]);
  if (!rc) rc = cress_download(x, CRESS_PRWTR, 0, prwtr);
  if (!rc) rc = cress_download(x, CRESS_PRICE, 0, price);
  cress_ctx_destroy(x);
  return rc;
}

// The first four calls of s_turbdrv (tubopt=2): every input and output
// array goes to the device, the kernels write the ranges the reference
// writes, and the outputs come back whole.
int cress_turbcoef(const cress_config *cfg, float dtb, int moist, const float *jcb, const float *jcb8u,
                   const float *jcb8v, const float *jcb8w, const float *pbr, const float *ptbr, const float *rbr,
                   const float *ppp, const float *ptpp, const float *qvp, const float *qallp, const float *up,
                   const float *vp, const float *wp, const float *tkep, float *t11, float *t22, float *t33,
                   float *t12, float *t13, float *t23, float *t31, float *t32, float *ssq, float *nsq8w, float *rkh,
                   float *rkv, float *priv) {
  const float *in[15] = {jcb, jcb8u, jcb8v, jcb8w, pbr, ptbr, rbr, ppp, ptpp, qvp, qallp, up, vp, wp, tkep};
  float *out[13] = {t11, t22, t33, t12, t13, t23, t31, t32, ssq, nsq8w, rkh, rkv, priv};
  for (const float *a : in)
    if (!a) return CRESS_EINVAL;
  for (float *a : out)
    if (!a) return CRESS_EINVAL;
  if (!cfg || cfg->ni < 4 || cfg->nj < 4 || cfg->nk < 4 || !(dtb > 0.0f) || !(cfg->dx > 0.0) ||
      !(cfg->dy > 0.0) || !(cfg->dz > 0.0) || cfg->cphopt < 0)
    return CRESS_EINVAL;
  const int sides[4] = {cfg->wbc, cfg->ebc, cfg->sbc, cfg->nbc};
  for (int s = 0; s < 4; s++)
    if (sides[s] != -1 && sides[s] != 7) return CRESS_EUNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf di[15], dout[13];
  int rc;
  for (int f = 0; f < 15; f++)
    if ((rc = to_dev(di[f], in[f], n))) return rc;
  for (int f = 0; f < 13; f++)
    if ((rc = to_dev(dout[f], out[f], n))) return rc;
  TurbArgs a = {};
  a.moist = moist ? 1 : 0;
  a.cphopt = cfg->cphopt;
  a.wbc = cfg->wbc, a.ebc = cfg->ebc, a.sbc = cfg->sbc, a.nbc = cfg->nbc, a.bbc = cfg->bbc, a.tbc = cfg->tbc;
  // the 4th-order reduction flags: 0 on periodic sides (sethalo.f90:536-580)
  a.iw = cfg->wbc == -1 ? 0 : 1, a.ie = cfg->ebc == -1 ? 0 : 1;
  a.js = cfg->sbc == -1 ? 0 : 1, a.jn = cfg->nbc == -1 ? 0 : 1;
  a.dx = (float)cfg->dx, a.dy = (float)cfg->dy, a.dz = (float)cfg->dz, a.dtb = dtb, a.thresq = (float)cfg->thresq;
  const float **ip[15] = {&a.jcb, &a.jcb8u, &a.jcb8v, &a.jcb8w, &a.pbr, &a.ptbr, &a.rbr, &a.pp,
                          &a.ptp, &a.qv, &a.qall, &a.u, &a.v, &a.w, &a.tke};
  for (int f = 0; f < 15; f++) *ip[f] = (const float *)di[f].p;
  float **op[13] = {&a.s11, &a.s22, &a.s33, &a.s12, &a.s13, &a.s23, &a.s31, &a.s32,
                    &a.ssq, &a.nsq8w, &a.rkh, &a.rkv, &a.priv};
  for (int f = 0; f < 13; f++) *op[f] = (float *)dout[f].p;
  turb_coef(g, a);
  HIPCHK(hipGetLastError());
  for (int f = 0; f < 13; f++) HIPCHK(hipMemcpy(out[f], dout[f].p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// s_turbdrv (tubopt=2) whole: the velocity forcings and the arrays it
// returns (rkh, rkv, priv, ssq, nsq8w, t13, t23, t33) go to the device and
// back; its work arrays t11, t22, t12, t31, t32 stay on the device.
int cress_turbdrv(const cress_config *cfg, float dtb, int moist, const float *jcb, const float *jcb8u,
                  const float *jcb8v, const float *jcb8w, const float *pbr, const float *ptbr, const float *rbr,
                  const float *up, const float *vp, const float *wp, const float *ppp, const float *ptpp,
                  const float *qvp, const float *tkep, const float *qallp, float *ufrc, float *vfrc, float *wfrc,
                  float *rkh, float *rkv, float *priv, float *ssq, float *nsq8w, float *t13, float *t23, float *t33) {
  const float *in[15] = {jcb, jcb8u, jcb8v, jcb8w, pbr, ptbr, rbr, ppp, ptpp, qvp, qallp, up, vp, wp, tkep};
  float *out[11] = {ufrc, vfrc, wfrc, rkh, rkv, priv, ssq, nsq8w, t13, t23, t33};
  for (const float *a : in)
    if (!a) return CRESS_EINVAL;
  for (float *a : out)
    if (!a) return CRESS_EINVAL;
  if (!cfg || cfg->ni < 4 || cfg->nj < 4 || cfg->nk < 4 || !(dtb > 0.0f) || !(cfg->dx > 0.0) ||
      !(cfg->dy > 0.0) || !(cfg->dz > 0.0) || cfg->cphopt < 0)
    return CRESS_EINVAL;
  const int sides[4] = {cfg->wbc, cfg->ebc, cfg->sbc, cfg->nbc};
  for (int s = 0; s < 4; s++)
    if (sides[s] != -1 && sides[s] != 7) return CRESS_EUNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf di[15], dout[11], dw[5];
  int rc;
  for (int f = 0; f < 15; f++)
    if ((rc = to_dev(di[f], in[f], n))) return rc;
  for (int f = 0; f < 11; f++)
    if ((rc = to_dev(dout[f], out[f], n))) return rc;
  for (int f = 0; f < 5; f++) {
    if ((rc = to_dev(dw[f], nullptr, n))) return rc;
    HIPCHK(hipMemset(dw[f].p, 0, n * sizeof(float)));
  }
  TurbArgs a = {};
  a.moist = moist ? 1 : 0;
  a.cphopt = cfg->cphopt;
  a.wbc = cfg->wbc, a.ebc = cfg->ebc, a.sbc = cfg->sbc, a.nbc = cfg->nbc, a.bbc = cfg->bbc, a.tbc = cfg->tbc;
  a.iw = cfg->wbc == -1 ? 0 : 1, a.ie = cfg->ebc == -1 ? 0 : 1;
  a.js = cfg->sbc == -1 ? 0 : 1, a.jn = cfg->nbc == -1 ? 0 : 1;
  a.dx = (float)cfg->dx, a.dy = (float)cfg->dy, a.dz = (float)cfg->dz, a.dtb = dtb, a.thresq = (float)cfg->thresq;
  const float **ip[15] = {&a.jcb, &a.jcb8u, &a.jcb8v, &a.jcb8w, &a.pbr, &a.ptbr, &a.rbr, &a.pp,
                          &a.ptp, &a.qv, &a.qall, &a.u, &a.v, &a.w, &a.tke};
  for (int f = 0; f < 15; f++) *ip[f] = (const float *)di[f].p;
  float **op[11] = {&a.ufrc, &a.vfrc, &a.wfrc, &a.rkh, &a.rkv, &a.priv, &a.ssq, &a.nsq8w, &a.s13, &a.s23, &a.s33};
  for (int f = 0; f < 11; f++) *op[f] = (float *)dout[f].p;
  float **wp5[5] = {&a.s11, &a.s22, &a.s12, &a.s31, &a.s32};
  for (int f = 0; f < 5; f++) *wp5[f] = (float *)dw[f].p;
  turb_drv(g, a);
  HIPCHK(hipGetLastError());
  for (int f = 0; f < 11; f++) HIPCHK(hipMemcpy(out[f], dout[f].p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// s_turbflx for one scalar: the fluxes come back over the ranges the
// reference writes, the rest of h1, h2, h3 as passed.
int cress_turbflx(const cress_config *cfg, const float *jcb, const float *s, const float *rkh8u,
                  const float *rkh8v, const float *rkv8w, float *h1, float *h2, float *h3) {
  const float *in[5] = {jcb, s, rkh8u, rkh8v, rkv8w};
  float *out[3] = {h1, h2, h3};
  for (const float *a : in)
    if (!a) return CRESS_EINVAL;
  for (float *a : out)
    if (!a) return CRESS_EINVAL;
  if (!cfg || cfg->ni < 4 || cfg->nj < 4 || cfg->nk < 3 || !(cfg->dx > 0.0) || !(cfg->dy > 0.0) || !(cfg->dz > 0.0))
    return CRESS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf di[5], dout[3];
  int rc;
  for (int f = 0; f < 5; f++)
    if ((rc = to_dev(di[f], in[f], n))) return rc;
  for (int f = 0; f < 3; f++)
    if ((rc = to_dev(dout[f], out[f], n))) return rc;
  turb_flx(g, (float)cfg->dx, (float)cfg->dy, (float)cfg->dz, (const float *)di[0].p, (const float *)di[1].p,
           (const float *)di[2].p, (const float *)di[3].p, (const float *)di[4].p, (float *)dout[0].p,
           (float *)dout[1].p, (float *)dout[2].p);
  HIPCHK(hipGetLastError());
  for (int f = 0; f < 3; f++) HIPCHK(hipMemcpy(out[f], dout[f].p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// s_turbs for one scalar: sfrc gains the flux divergence over its interior.
int cress_turbs(const cress_config *cfg, const float *jcb8u, const float *jcb8v, const float *h1, const float *h2,
                const float *h3, float *sfrc) {
  const float *in[5] = {jcb8u, jcb8v, h1, h2, h3};
  for (const float *a : in)
    if (!a) return CRESS_EINVAL;
  if (!sfrc || !cfg || cfg->ni < 4 || cfg->nj < 4 || cfg->nk < 4 || !(cfg->dx > 0.0) || !(cfg->dy > 0.0) ||
      !(cfg->dz > 0.0))
    return CRESS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf di[5], d;
  int rc;
  for (int f = 0; f < 5; f++)
    if ((rc = to_dev(di[f], in[f], n))) return rc;
  if ((rc = to_dev(d, sfrc, n))) return rc;
  turb_s(g, (float)cfg->dx, (float)cfg->dy, (float)cfg->dz, (const float *)di[0].p, (const float *)di[1].p,
         (const float *)di[2].p, (const float *)di[3].p, (const float *)di[4].p, (float *)d.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(sfrc, d.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// s_kh8uv: rkh8u and rkh8v come back over the ranges the reference writes.
int cress_kh8uv(const cress_config *cfg, const float *rkh, float *rkh8u, float *rkh8v) {
  if (!cfg || !rkh || !rkh8u || !rkh8v || cfg->ni < 3 || cfg->nj < 3 || cfg->nk < 2) return CRESS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf a, b, d;
  int rc;
  if ((rc = to_dev(a, rkh, n)) || (rc = to_dev(b, rkh8u, n)) || (rc = to_dev(d, rkh8v, n))) return rc;
  turb_kh8uv(g, (const float *)a.p, (float *)b.p, (float *)d.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(rkh8u, b.p, n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rkh8v, d.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// s_advp alone is the pressure forcing kernel without sponge or smoothing
// (its forcing is the advection term only, written over the interior).
int cress_advp(const cress_config *cfg, const float *jcb8u, const float *jcb8v, const float *jcb8w,
               const float *u, const float *v, const float *wc, const float *pp, float *pfrc) {
  if (!cfg || !jcb8u || !jcb8v || !jcb8w || !u || !v || !wc || !pp || !pfrc) return CRESS_EINVAL;
  if (cfg->ni < 7 || cfg->nj < 7 || cfg->nk < 7 || cfg->advopt < 1 || cfg->advopt > 3) return CRESS_EINVAL;
  const int sides[4] = {cfg->wbc, cfg->ebc, cfg->sbc, cfg->nbc};
  for (int s = 0; s < 4; s++)
    if (sides[s] != -1 && sides[s] != 7) return CRESS_EUNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  Consts c;
  fold_consts(cfg, &c);
  Opts o = {};
  o.advopt = cfg->advopt;
  o.iw = cfg->wbc == -1 ? 0 : 1;
  o.ie = cfg->ebc == -1 ? 0 : 1;
  o.js = cfg->sbc == -1 ? 0 : 1;
  o.jn = cfg->nbc == -1 ? 0 : 1;
  DevBuf a, b, d, e, f, h, p, q;
  int rc;
  if ((rc = to_dev(a, jcb8u, n)) || (rc = to_dev(b, jcb8v, n)) || (rc = to_dev(d, jcb8w, n)) ||
      (rc = to_dev(e, u, n)) || (rc = to_dev(f, v, n)) || (rc = to_dev(h, wc, n)) || (rc = to_dev(p, pp, n)) ||
      (rc = to_dev(q, pfrc, n)))
    return rc;
  const LaunchTune tn = tune_default();
  const Tiles tp = tiles(g.ni - 1, g.nj - 3, 64, tn.ff_by, g.nk - 3, tn.ff_order);
  const float *pj = (const float *)a.p, *pp_ = (const float *)p.p;
  hipLaunchKernelGGL(k_force_p<false>, flat(tp), dim3(64, tn.ff_by), 0, 0, tp, g, o, c, pj, pj,
                     (const float *)b.p, (const float *)d.p, (const float *)e.p, (const float *)f.p,
                     (const float *)h.p, pp_, pp_, pj, (float *)q.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(pfrc, q.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

// The three staggered-point boundary fills share one kernel; kind 0/1/2 =
// bc8u/bc8v/bc8w with the sides' options lo, hi.
static int bc8(const cress_config *cfg, int kind, int lo, int hi, int kmax, float *var) {
  if (!cfg || !var || cfg->ni < 4 || cfg->nj < 4 || cfg->nk < 4) return CRESS_EINVAL;
  if (kind < 2 && (kmax < 1 || kmax > cfg->nk)) return CRESS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)(g.ni + 2) * (g.nj + 2) * (size_t)(kind == 2 ? g.nk : kmax);
  DevBuf a;
  int rc;
  if ((rc = to_dev(a, var, n))) return rc;
  const long pts = kind == 2 ? (long)(g.ni + 2) * (g.nj + 2) : (long)(kind == 0 ? g.nj + 2 : g.ni + 2) * kmax;
  hipLaunchKernelGGL(k_bc8, dim3((unsigned)((pts + 63) / 64), 2), dim3(64), 0, 0, g, kind, lo, hi, kmax,
                     (float *)a.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(var, a.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}

int cress_bc8u(const cress_config *cfg, int kmax, float *var8u) { return cfg ? bc8(cfg, 0, cfg->wbc, cfg->ebc, kmax, var8u) : CRESS_EINVAL; }
int cress_bc8v(const cress_config *cfg, int kmax, float *var8v) { return cfg ? bc8(cfg, 1, cfg->sbc, cfg->nbc, kmax, var8v) : CRESS_EINVAL; }
int cress_bc8w(const cress_config *cfg, float *var8w) { return cfg ? bc8(cfg, 2, cfg->bbc, cfg->tbc, cfg->nk, var8w) : CRESS_EINVAL; }

int cress_bcycle(const cress_config *cfg, int iwsnd, int iwrcv, int iesnd, int iercv, int jssnd, int jsrcv,
                 int jnsnd, int jnrcv, int kmax, float *var) {
  if (!cfg || !var || kmax < 1 || kmax > cfg->nk) return CRESS_EINVAL;
  const int ni = cfg->ni, nj = cfg->nj;
  const int is[4] = {iwsnd, iwrcv, iesnd, iercv}, js[4] = {jssnd, jsrcv, jnsnd, jnrcv};
  for (int t = 0; t < 4; t++)
    if (is[t] < 0 || is[t] > ni + 1 || js[t] < 0 || js[t] > nj + 1) return CRESS_EINVAL;
  if (!(cfg->wbc == -1 && cfg->ebc == -1 && cfg->sbc == -1 && cfg->nbc == -1)) return CRESS_EUNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(ni, nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf a;
  int rc = to_dev(a, var, n);
  if (rc) return rc;
  FillSet fs;
  fs.n = 1;
  fill_spec(&fs.f[0], (float *)a.p, iwsnd, iwrcv, iesnd, iercv, jssnd, jsrcv, jnsnd, jnrcv, kmax);
  hipLaunchKernelGGL(k_bcycle_x, dim3((g.nj + 2 + 63) / 64, kmax, 1), dim3(64), 0, 0, g, fs);
  hipLaunchKernelGGL(k_bcycle_y, dim3((g.ni + 2 + 63) / 64, kmax, 1), dim3(64), 0, 0, g, fs);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(var, a.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}


// Attach a network to a context (mode 1: record pairs from the plain
// scheme into its capture buffers; 2: use it for the source terms of
// every active cell inside its envelope) or detach it (mlp NULL or mode 0).
int cress_ctx_mlp_cells(cress_ctx *x, long *cells) {
  if (!x || !cells) return CRESS_EINVAL;
  *cells = 0;
  if (x->mlp_mode != 2 || !x->ml.count) return CRESS_OK;
  unsigned c = 0;
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(&c, x->ml.count, sizeof c, hipMemcpyDeviceToHost));
  *cells = (long)c;
  return CRESS_OK;
}

int cress_ctx_set_mlp(cress_ctx *x, cress_mlp *mlp, int mode) {
  if (!x || mode < 0 || mode > 2 || (mode && !mlp)) return CRESS_EINVAL;
  if (mode && x->cfg.cphopt != 2) return CRESS_EUNSUPPORTED;
  if (mode == 2 && !x->ml.flag) {
    const long n = x->n;
    float *f = nullptr;
    int rc = dalloc(x, &f, (size_t)n);
    if (!rc) x->ml.flag = (int *)f;
    if (!rc) rc = dalloc(x, &f, (size_t)n);
    if (!rc) x->ml.list = (int *)f;
    if (!rc) rc = dalloc(x, &f, (size_t)n * MLP_INP / 2);
    if (!rc) x->ml.x = (unsigned short *)f;
    if (!rc) rc = dalloc(x, &f, 1);
    if (!rc) x->ml.count = (unsigned *)f;
    if (!rc) rc = dalloc(x, &f, (size_t)n * MLP_NOUT);
    if (!rc) x->ml.out = f;
    if (rc) {
      x->ml = MlpStep{nullptr, nullptr, nullptr, nullptr, nullptr};
      return rc;
    }
  }
  x->mlp = mode ? mlp : nullptr;
  x->mlp_mode = mode;
  return CRESS_OK;
}

// s_forcetke: the advection of the present TKE by the kernel s_advs uses in
// the step, then the closure's terms (turb_tke).  rkh, rkv, tkefrc and rkv8s
// go to the device and back whole.
int cress_forcetke(const cress_config *cfg, const float *jcb, const float *jcb8u, const float *jcb8v,
                   const float *jcb8w, const float *rst, const float *rstxu, const float *rstxv, const float *rstxwc,
                   const float *tke, const float *tkep, const float *rbct, const float *priv, const float *nsq8w,
                   const float *ssq, float *rkh, float *rkv, float *tkefrc, float *rkv8s) {
  const float *in[14] = {jcb, jcb8u, jcb8v, jcb8w, rst, rstxu, rstxv, rstxwc, tke, tkep, priv, nsq8w, ssq, rbct};
  float *out[4] = {rkh, rkv, tkefrc, rkv8s};
  if (!cfg) return CRESS_EINVAL;
  const int vsp = cfg->vspopt >= 1 && cfg->vspvar[9] == 'o';
  for (int f = 0; f < 14; f++)
    if (!in[f] && (f < 13 || vsp)) return CRESS_EINVAL;
  for (float *a : out)
    if (!a) return CRESS_EINVAL;
  if (cfg->ni < 7 || cfg->nj < 7 || cfg->nk < 7 || cfg->advopt < 1 || cfg->advopt > 3 || !(cfg->dx > 0.0) ||
      !(cfg->dy > 0.0) || !(cfg->dz > 0.0) || (vsp && (cfg->ksp0[1] < 2 || cfg->ksp0[1] > cfg->nk)))
    return CRESS_EINVAL;
  if (cfg->smtopt % 10 != 0) return CRESS_EUNSUPPORTED;
  const int sides[4] = {cfg->wbc, cfg->ebc, cfg->sbc, cfg->nbc};
  for (int s = 0; s < 4; s++)
    if (sides[s] != -1 && sides[s] != 7) return CRESS_EUNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CRESS_ENODEV;
  const Grid g = make_grid(cfg->ni, cfg->nj, cfg->nk);
  const size_t n = (size_t)g.n();
  DevBuf di[14], dout[4];
  int rc;
  for (int f = 0; f < 13; f++)
    if ((rc = to_dev(di[f], in[f], n))) return rc;
  if (vsp && (rc = to_dev(di[13], rbct, 2 * (size_t)cfg->ni * cfg->nj * cfg->nk))) return rc;
  for (int f = 0; f < 4; f++)
    if ((rc = to_dev(dout[f], out[f], n))) return rc;
  Consts c;
  fold_consts(cfg, &c);
  Opts o = {};
  o.advopt = cfg->advopt;
  o.iw = cfg->wbc == -1 ? 0 : 1;
  o.ie = cfg->ebc == -1 ? 0 : 1;
  o.js = cfg->sbc == -1 ? 0 : 1;
  o.jn = cfg->nbc == -1 ? 0 : 1;
  ScalarIn si = {};
  si.s[0] = (const float *)di[8].p;
  si.frc[0] = (float *)dout[2].p;
  si.rstxu = (const float *)di[5].p;
  si.rstxv = (const float *)di[6].p;
  si.rstxwc = (const float *)di[7].p;
  hipLaunchKernelGGL(k_advs_only, grid2(g.ni - 3, g.nj - 3, 64, 4, g.nk - 3), dim3(64, 4), 0, 0, g, o, c, si);
  TkeArgs a = {};
  a.dx = (float)cfg->dx, a.dy = (float)cfg->dy, a.dz = (float)cfg->dz;
  a.vsp = vsp, a.ksp0 = cfg->ksp0[1];
  a.jcb = (const float *)di[0].p, a.jcb8u = (const float *)di[1].p, a.jcb8v = (const float *)di[2].p;
  a.jcb8w = (const float *)di[3].p, a.rst = (const float *)di[4].p, a.tkep = (const float *)di[9].p;
  a.priv = (const float *)di[10].p, a.nsq8w = (const float *)di[11].p, a.ssq = (const float *)di[12].p;
  a.rbct = vsp ? (const float *)di[13].p : nullptr;
  a.rkh = (float *)dout[0].p, a.rkv = (float *)dout[1].p, a.tkefrc = (float *)dout[2].p, a.rkv8s = (float *)dout[3].p;
  turb_tke(g, a);
  HIPCHK(hipGetLastError());
  for (int f = 0; f < 4; f++) HIPCHK(hipMemcpy(out[f], dout[f].p, n * sizeof(float), hipMemcpyDeviceToHost));
  return CRESS_OK;
}
