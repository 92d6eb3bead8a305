/* cress_hip.h -- C ABI of the MI355X (gfx950) implementation of one CReSS
 * large time step.
 *
 * Two tiers:
 *
 *  1. Device-resident solver context (the performance path).  The host
 *     uploads the base state and the three time levels once, then calls
 *     cress_large_step() for as many steps as it likes; nothing crosses PCIe
 *     between steps.  This replaces the reference's time-loop body
 *       s_intgdrv   (SrcOrig/intgdrv.f90:111-1133, called from slvdrv.f90:1306)
 *     with fdmdrv/forcedrv/smalldrv/hevi/steps/cloudphy/timeflt/satadjst/
 *     shift2nd/swp2nxt inside it, for the plain-path options.
 *
 *  2. Per-routine drop-in shims with the reference's argument lists and
 *     explicit-shape host arrays real(0:ni+1,0:nj+1,1:nk) (column-major,
 *     x fastest).  Each copies its arrays to the device, runs the same HIP
 *     kernel the context uses, and copies the outputs back.
 *
 * Conventions: every function returns 0 on success or a negative CRESS_E*
 * code; cress_error_string() names it.  Option values are passed already
 * resolved (the caller reads the namelist; the library never does).  One
 * host thread per context; each context owns one HIP stream.  All arrays are
 * IEEE binary32 unless stated.
 */
#ifndef CRESS_HIP_H
#define CRESS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CRESS_OK = 0,
  CRESS_EINVAL = -1,      /* bad argument / unsupported option combination */
  CRESS_ENODEV = -2,      /* no usable GPU */
  CRESS_EHIP = -3,        /* HIP runtime error */
  CRESS_ENOMEM = -4,      /* device allocation failed */
  CRESS_ERANGE = -5,      /* field id / level out of range */
  CRESS_EUNSUPPORTED = -6, /* option outside the implemented plain path */
  CRESS_ETOOBIG = -7       /* grid larger than a fixed-size buffer allows */
};

/* Resolved run options.  Names and meanings follow the reference namelist
 * (Doc/0rig/readme_user.conf.txt); dims follow setdim.f90:196-199
 * (ni = nx+3, nj = ny+3, nk = nz+3). */
typedef struct cress_config {
  int ni, nj, nk;
  int wbc, ebc, sbc, nbc;   /* lateral BCs after the one-rank remap: -1 periodic
                               (chknlslv.f90:751-790) on all four sides, or 7
                               (open, constant phase speed) on all four */
  int bbc, tbc;             /* bottom/top: 2 rigid, 3 zero gradient */
  int advopt;               /* 1,2: 2nd/4th order horizontal; 3: 4th order 3-D */
  int impopt;               /* 1: implicit vertical (Thomas, gaussel) */
  int gwmopt;               /* 0 */
  int buyopt;               /* 0/1: buoyancy in the implicit solve */
  int divopt;               /* 0/1: divergence damping */
  int cphopt;               /* 0 dry-ish (no microphysics), 2 cold-rain bulk */
  int haiopt;               /* 0 */
  int vspopt;               /* 0 off, 2 Rayleigh sponge to the base state */
  int coropt;               /* 0/1 Coriolis on u,v */
  int sthopt;               /* 0 uniform dz, >=1 stretched (wc = w/jcb8w) */
  int ksp0[2];              /* lowest sponge level (vspdmp.f90:167-232) */
  char vspvar[10];          /* 'o'/'x' for u,v,w,p,pt,qv,q,... */
  int nsstp;                /* acoustic substeps per large step (slvstep.f90:176) */
  double dx, dy, dz;        /* m; dz is the computational (zeta) spacing */
  double dtbig, dtsml;      /* s */
  double weicoe, filcoe, thresq;
  double gwave;             /* open sides (wbc..nbc = 7): the constant phase
                               speed of the radiative fills, m/s
                               (&boundry gwave, phvuvw.f90:334-338) */
  int smtopt;               /* 0 none; 2, 3: 4th-order smoothing of the past
                               level, horizontal / also vertical (&smoother) */
  double smhcoe, smvcoe;    /* smoothing coefficients, 1/s */
} cress_config;

/* Field ids for upload/download.  Base state and metrics have one level. */
enum {
  /* base state / metrics: level 0 only */
  CRESS_JCB = 0, CRESS_JCB8U, CRESS_JCB8V, CRESS_JCB8W,
  CRESS_UBR, CRESS_VBR, CRESS_PBR, CRESS_PTBR, CRESS_QVBR, CRESS_RBR,
  CRESS_RST, CRESS_RST8U, CRESS_RST8V, CRESS_RST8W, CRESS_RCSQ,
  CRESS_NBASE,
  /* prognostics: level 0 past, 1 present, 2 future */
  CRESS_U = 32, CRESS_V, CRESS_W, CRESS_PP, CRESS_PTP,
  CRESS_QV, CRESS_QC, CRESS_QR, CRESS_QI, CRESS_QS, CRESS_QG, CRESS_NCI,
  CRESS_NPROG_END,
  /* 2-D / special layouts */
  CRESS_FC = 64,            /* (0:ni+1,0:nj+1,1:2) Coriolis, component 1 used */
  CRESS_RBCT = 65,          /* (1:ni,1:nj,1:nk,1:2) sponge coefficients */
  CRESS_PRWTR = 66,         /* (0:ni+1,0:nj+1,1:2,1:2) rain/cloud precip rate, accum */
  CRESS_PRICE = 67          /* (0:ni+1,0:nj+1,1:2,1:3) ice/snow/graupel precip */
};

typedef struct cress_ctx cress_ctx;

const char *cress_error_string(int code);
int cress_device_count(int *n);

/* ---- Tier 1: device-resident context --------------------------------- */
int cress_ctx_create(const cress_config *cfg, cress_ctx **out);
int cress_ctx_destroy(cress_ctx *ctx);
/* Number of floats of one field (its layout as documented above). */
int cress_field_size(const cress_ctx *ctx, int field, size_t *nfloat);
int cress_upload(cress_ctx *ctx, int field, int level, const float *host);
int cress_download(cress_ctx *ctx, int field, int level, float *host);
/* Device layout of the 3-D fields: element (i,j,k) of a field lies
 * i + row*j + plane*(k-1) floats from its device address, row >= ni+2
 * (rows are padded to 128-byte multiples); upload/download convert from and
 * to the host layout row = ni+2.  2-D fields (fc, rbct, prwtr, price) are
 * stored as on the host. */
int cress_field_pitch(const cress_ctx *ctx, size_t *row, size_t *plane);
/* Device address of a field (for callers that fill device memory
 * themselves, in the layout of cress_field_pitch); valid until the next
 * cress_large_step().  A base-state or
 * metric field uploaded with every column equal (flat terrain, a
 * horizontally uniform base state) is also kept as one value per level and
 * read from there; asking for its device address drops that copy, since the
 * caller may write the field. */
int cress_field_device_ptr(cress_ctx *ctx, int field, int level, void **dptr);
/* Advance nsteps large steps.  first != 0 makes the first of them the
 * reference's ctime=0 step: forward step, dtb=dtbig/2, dts=dtsml/2, no
 * Asselin filter (timeint.f90:160-163, intgdrv.f90:968). */
int cress_large_step(cress_ctx *ctx, int nsteps, int first);
int cress_sync(cress_ctx *ctx);
/* Device time of the last cress_large_step call's kernels, per stage, in ms
 * (measured with HIP events on the context's stream when timing is on).
 * on: 0 off; 1 every kernel family and the stages (the host waits for each
 * step's events); 2 + id: only the launches of kernel family id are
 * bracketed, nothing else is recorded and the steps run back to back. */
int cress_set_timing(cress_ctx *ctx, int on);
int cress_stage_times(cress_ctx *ctx, double *ms, int nstage);
/* Per-kernel-family device time (HIP events around every launch while
 * timing is on), accumulated since the last cress_reset_timing(). */
int cress_kernel_count(void);
const char *cress_kernel_name(int id);
int cress_kernel_times(cress_ctx *ctx, double *total_ms, long *launches, int n);
int cress_reset_timing(cress_ctx *ctx);

/* Base state and metrics from a 1-D sounding, computed in binary64 on the
 * device and rounded once into the binary32 fields (flat terrain, every
 * column the same): the reference's setup chain s_stretch
 * (stretch.f90:174-359, sthopt 0/1 from the context's options), s_jacobian
 * (jacobian.f90:243-251), s_intrpsnd (intrpsnd.f90:99-301: vint13 linear
 * interpolation of the sounding to the scalar heights), s_setbase and
 * s_bcbase (setbase.f90:171-185, bcbase.f90:120-190), s_baserho
 * (baserho.f90:153-170), s_var8uvw with bc8w (var8uvw.f90:166-294),
 * s_sndwave (sndwave.f90:119) and s_vspdmp (vspdmp.f90:160-285, vspopt=2,
 * which also sets the sponge's lowest level ksp0).  Replaces every base
 * field (CRESS_JCB..CRESS_RCSQ), rbct and fc.  The top slot nk of the
 * scalar fields and of jcb, which the reference leaves unset, repeats nk-1. */
typedef struct cress_sounding {
  int nlev;                 /* the sounding has levels 0..nlev (nlev >= 1) */
  const double *z;          /* heights, m, increasing */
  const double *u, *v;      /* wind, m/s */
  const double *p;          /* pressure, Pa */
  const double *pt;         /* potential temperature, K */
  const double *qv;         /* water vapour mixing ratio, kg/kg */
  double zsfc, dzmin, layer1, layer2;  /* &gridsth: surface height and the
                                          stretching's layers (sthopt=1) */
  double botbar, vspbar;    /* sponge above botbar (m), coefficient (1/s) */
  double fc;                /* Coriolis parameter (constant f-plane), 1/s */
} cress_sounding;
int cress_set_base(cress_ctx *ctx, const cress_sounding *snd);
/* The setup's binary64 per-level columns (k = 0..nk, slot 0 unused), ncol
 * of them in this order: zph (zeta levels), zph8s, ubr, vbr, pbr, ptbr, qvbr,
 * rbr, jcb, jcb8w, rst, rst8w, rcsq, rbct (vspopt=2), ptvbr, pibr. */
#define CRESS_NSETUPCOL 16
int cress_base_columns(const cress_ctx *ctx, double *out, int ncol);

/* Launch shapes of the tunable kernel families, as 9 ints: fused
 * ptil/stepuv rows per workgroup (4, 8, 16), halo layout (0 one point per
 * thread, 1 slots), tile order (1 x-fastest, 2 z-fastest); column kernel
 * rows (1, 2, 4) and LDS bytes per workgroup (caps occupancy); scalar
 * forcing rows and order; u/v, w, p' forcing rows and order.  No
 * counterpart in the reference (tile shapes are a GPU matter).  Results do
 * not depend on the shapes.  cress_autotune times every candidate on the
 * context's own grid and keeps the fastest (cached per process by grid and
 * options); it writes only scratch arrays and the future level, so it may
 * run between any two large steps, with timing off.  shape may be NULL. */
#define CRESS_NLAUNCH 9
int cress_autotune(cress_ctx *ctx, int *shape, int n);
int cress_set_launch(cress_ctx *ctx, const int *shape, int n);
int cress_get_launch(const cress_ctx *ctx, int *shape, int n);

/* Mass and energy budgets of one level (0 past, 1 present, 2 future),
 * reduced over the interior (i,j,k = 2..n-2) in binary64 on the device, in a
 * fixed order (repeatable bit for bit).  The reference monitors extremes
 * (getmxn.f90:173-232) and accumulates the surface precipitation
 * (upwqp.f90:199-202); out gets CRESS_NDIAG values:
 *   0 vapour mass, sum rst qv dV (kg; rst = rho_bar J, dV = dx dy dz)
 *   1 liquid water mass (qc + qr)          2 ice mass (qi + qs + qg)
 *   3 accumulated surface precipitation over the interior columns, as
 *     water mass (kg): sum of the accumulations (m) x rhow x dx dy
 *   4 dry-air mass, sum J rho_d dV with rho_d = p eps/((eps + qv) Rd T)
 *   5 kinetic energy, sum J rho_d (u^2+v^2+w^2)/2 dV (velocities averaged to
 *     the scalar point)            6 internal energy, sum J rho_d cv T dV
 *   7 potential energy, sum J rho_d g z dV (z of the scalar levels over flat
 *     terrain; NaN when jcb varies horizontally)
 *   8 sum rst theta' dV            9, 10 max and min w
 *   11, 12 max and min theta' */
#define CRESS_NDIAG 13
int cress_diag(cress_ctx *ctx, int level, double *out);

/* ---- Ensemble: contexts of any sizes stepped together ------------------
 * The reference runs one model per job; these have no counterpart there.
 * They batch independent contexts (members) on one GPU: a cycle of large
 * steps of every member (3 steps, 6 when a member's nsstp is odd) is
 * captured once into one hipGraph whose parallel branches are the members'
 * own streams, and each cycle is then one graph launch.  Results are
 * bitwise those of stepping each member with cress_large_step.  Members
 * keep their own state and are uploaded/downloaded as usual; timing must be
 * off while they are stepped here.  cress_ens_step's first != 0 makes the
 * first step every member's cold-start step (run without the graph). */
typedef struct cress_ens cress_ens;
int cress_ens_create(cress_ctx *const *members, int n, cress_ens **out);
int cress_ens_step(cress_ens *e, int nsteps, int first);
int cress_ens_sync(cress_ens *e);
int cress_ens_destroy(cress_ens *e);

/* ---- Tier 2: per-routine drop-in shims (host arrays) -------------------
 * Each mirrors the reference subroutine named in its comment: same array
 * arguments in the same order, options taken from cfg instead of fp ids. */

/* s_gaussel(fpimpopt,istr,iend,jstr,jend,kstr,kend,ni,nj,kmax,rr,ss,tt,ff,pv)
 * SrcOrig/gaussel.f90:72-73, impopt=1 path.  ss and ff are overwritten. */
int cress_gaussel(int istr, int iend, int jstr, int jend, int kstr, int kend,
                  int ni, int nj, int nk, float *rr, float *ss, float *tt, float *ff);

/* cress_gaussel without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * loops over many small column boxes, callers that forbid hipMalloc).  The
 * four arrays are staged in a fixed device array of
 * CRESS_GAUSSEL_SCRATCH_FLOATS floats per GPU (held until the process
 * ends), so (ni+2)(nj+2)nk may be at most a quarter of that, else
 * CRESS_ETOOBIG.  Calls that overlap on one GPU share the array and race.
 * Same arguments, results and errors as cress_gaussel otherwise. */
#define CRESS_GAUSSEL_SCRATCH_FLOATS (8 * 1024 * 1024)
int cress_gaussel_noalloc(int istr, int iend, int jstr, int jend, int kstr, int kend,
                          int ni, int nj, int nk, float *rr, float *ss, float *tt, float *ff);

/* s_advs(fpadvopt,...,ni,nj,nk,rstxu,rstxv,rstxwc,s,sfrc,vadv,tmp1,tmp2,tmp3)
 * SrcOrig/advs.f90:75-77.  Writes sfrc over the interior. */
int cress_advs(const cress_config *cfg, const float *rstxu, const float *rstxv,
               const float *rstxwc, const float *s, float *sfrc);

/* s_advuvw(fpadvopt,fpiwest,...,fpdziv,ni,nj,nk,rstxu,rstxv,rstxwc,u,v,w,
 *          ufrc,vfrc,wfrc,hadv,vadv,tmp1,tmp2,tmp3)  SrcOrig/advuvw.f90:75-78.
 * Adds the advection of u, v, w to ufrc, vfrc, wfrc over the reference's
 * ranges; the lateral sides are periodic (-1) or open (7), which decides
 * where the 4th-order terms stop (iwest..jnorth). */
int cress_advuvw(const cress_config *cfg, const float *rstxu, const float *rstxv, const float *rstxwc,
                 const float *u, const float *v, const float *w, float *ufrc, float *vfrc, float *wfrc);

/* s_hevi(fpgwmopt,fpbuyopt,isstp,dts,gtinc,ni,nj,nk,j31,j32,jcb,jcb8u,jcb8v,
 *        jcb8w,mf,mf8u,mf8v,rmf,rmf8u,rmf8v,ubr,vbr,ptbr,rbr,rst,rst8u,rst8v,
 *        rst8w,rcsq,ufrc,vfrc,wfrc,pfrc,ptfrc,ucpx,...,ptptd,uf,vf,wf,wc,ppf,
 *        ptpf,tmp1..tmp6)  SrcOrig/hevi.f90:88-95, for gwmopt=0, flat terrain
 * (j31=j32=0), no map factors and periodic sides: one acoustic substep of
 * uf, vf, wf, wc, ppf in place, with the halo fills of stepuv, stepwi and
 * steppi.  buyopt and divopt come from cfg. */
int cress_hevi(const cress_config *cfg, float dts, const float *jcb, const float *jcb8u, const float *jcb8v,
               const float *jcb8w, const float *rbr, const float *rst, const float *rst8u, const float *rst8v,
               const float *rst8w, const float *rcsq, const float *ufrc, const float *vfrc, const float *wfrc,
               const float *pfrc, float *uf, float *vf, float *wf, float *wc, float *ppf);

/* s_steps(fpgpvvar,...,fmois,dtb,gtinc,atinc,ni,nj,nk,nqw,nnw,nqi,nni,nqa,
 *         qvbr,rst,ptp,ptpp,qv,qvp,qwtr,qwtrp,...,ptfrc,qvfrc,qwtrf,...,
 *         ptpf,qvf,dtdrst)  SrcOrig/steps.f90:114-127, for gwmopt=0,
 * advopt<=3, periodic sides, no nudging, aerosol, tracer or TKE: the
 * leapfrog of theta', qv and, with cphopt=2, qc, qr (qwtr*(..,1:2)), qi,
 * qs, qg (qice*(..,1:3)) and nci, then the halo fills and vbcs.  qwtrf,
 * qicef, nicef hold the species forcings on entry and the future level on
 * return, as in the reference; ptpf, qvf and dtdrst (= 2 dtb / rst over
 * the interior) are written.  The species arrays are ignored for
 * cphopt=0. */
int cress_steps(const cress_config *cfg, float dtb, const float *rst, const float *ptpp, const float *ptfrc,
                const float *qvp, const float *qvfrc, const float *qwtrp, const float *qicep, const float *nicep,
                float *qwtrf, float *qicef, float *nicef, float *ptpf, float *qvf, float *dtdrst);

/* The TKE lines of s_steps (SrcOrig/steps.f90:997-1012, 4170-4172,
 * 4350-4352, 4493-4495) for tubopt>=2, periodic sides: tkef holds the TKE
 * forcing on entry and the future TKE on return, max(0, tkep + tkef *
 * 2 dtb / rst) over the interior, then the halo fills and vbcs as for the
 * other scalars; bitwise.  With cress_steps it covers s_steps with TKE; it
 * has no separate counterpart in the reference. */
int cress_stepstke(const cress_config *cfg, float dtb, const float *rst, const float *tkep, float *tkef);

/* s_cloudphy(fpadvopt,fpcphopt,fphaiopt,nclstp,dtb,dtcl,ni,nj,nk,nqw,nnw,
 *            nqi,nni,km,jcb,jcb8w,pbr,ptbr,rbr,rst,wf,ppp,ptpp,qvp,qwtrp,
 *            qicep,qallp,ptpf,qvf,qwtrf,nwtrf,nwtrp,qicef,nicef,nicep,
 *            qcwtrf,qcicef,prwtr,price,rbv,pi,p,tmp1..tmp11)
 * SrcOrig/cloudphy.f90:84-91, for the cold-rain scheme (cphopt=2,
 * haiopt=0, advopt<=3): getvdens, getexner, diagnw, diagnsg, adjstnci,
 * termblk, srcblk, fallblk on the future level (ptpf, qvf, qwtrf(..,1:2),
 * qicef(..,1:3), nicef) from the past level, over 2 dtb; prwtr
 * (0:ni+1,0:nj+1,1:2,1:2) and price (..,1:2,1:3) get the surface rate and
 * accumulation.  Within 1e-5 of the FP32 reference (binary32 exp/log). */
int cress_cloudphy(const cress_config *cfg, float dtb, const float *jcb, const float *pbr, const float *ptbr,
                   const float *rbr, const float *rst, const float *ppp, const float *ptpp, const float *qvp,
                   const float *qwtrp, const float *qicep, const float *nicep, float *ptpf, float *qvf,
                   float *qwtrf, float *qicef, float *nicef, float *prwtr, float *price);

/* ---- Learned microphysics (no counterpart in the reference) ----
 * A 9 -> 128 -> 128 -> 128 -> 7 GELU network standing in for the local
 * source terms of s_srcblk (srcblk.f90:488-663, called from
 * coldblk.f90:354-358; sedimentation and saturation adjustment stay on the
 * plain path).  Features (raw units): T, p, rho, qv, qc, qr, qi, qs, qg of
 * the past level; targets: the change srcblk makes to the future level per
 * second of the microphysics step, as dT/dt (= pi dtheta'/dt) and dqv/dt ..
 * dqg/dt.  BF16 weights and activations into the matrix cores, FP32
 * accumulation and master weights, Adam.  Parameters are the FP32 block
 * W1[16][128] b1 W2[128][128] b2 W3[128][128] b3 W4[128][16] b4[16]
 * (cress_mlp_nparam() floats; rows/columns beyond 9 inputs / 7 outputs are
 * zero); the normalisation block (cress_mlp_nnorm() floats) is in_mu[16]
 * in_rsd[16] out_mu[16] out_sd[16] env_lo[16] env_hi[16]. */
typedef struct cress_mlp cress_mlp;
int cress_mlp_nparam(void);
int cress_mlp_nnorm(void);
/* capacity: training pairs held on the device; max_batch: largest batch. */
int cress_mlp_create(long capacity, int max_batch, cress_mlp **out);
int cress_mlp_destroy(cress_mlp *m);
int cress_mlp_init(cress_mlp *m, unsigned long long seed);       /* Glorot-uniform weights */
int cress_mlp_set(cress_mlp *m, const float *params, const float *norm);   /* either may be NULL */
int cress_mlp_get(cress_mlp *m, float *params, float *norm);
int cress_mlp_samples(cress_mlp *m, long *count);
int cress_mlp_clear_samples(cress_mlp *m);
int cress_mlp_put_samples(cress_mlp *m, const float *in, const float *tgt, long n);   /* rows of 9 / 7 */
int cress_mlp_get_samples(cress_mlp *m, long first, long n, float *in, float *tgt);
/* normalisation and the envelope [min - margin*range, max + margin*range]
 * from the held pairs */
int cress_mlp_fit_norm(cress_mlp *m, float margin);
/* the step's fused inference kernel on n host rows of raw features */
int cress_mlp_forward(cress_mlp *m, const float *in, long n, float *out);
/* the training kernels' gradient of the mean squared error (normalised
 * targets) on n host rows, no update */
int cress_mlp_grad(cress_mlp *m, const float *in, const float *tgt, int n, float *grad, double *loss);
/* steps Adam steps on random batches of the held pairs; losses[steps] may be NULL */
int cress_mlp_train(cress_mlp *m, int steps, int batch, float lr, unsigned seed, double *losses);
/* mode 0: srcblk; 1: srcblk, recording (features, targets) of every active
 * cell into m; 2: the network for active cells inside its envelope, srcblk
 * for the other active cells, nothing where no process is active. */
int cress_ctx_set_mlp(cress_ctx *x, cress_mlp *m, int mode);
/* Cells the network handled in the last large step (mode 2), else 0. */
int cress_ctx_mlp_cells(cress_ctx *x, long *cells);

/* The first four calls of s_turbdrv (SrcOrig/turbdrv.f90:245-276) with
 * tubopt=2 (1.5-order TKE closure), isoopt=1, trnopt=0, mfcopt=0, sfcopt=0:
 *   s_bruntv(fpcphopt,fpdziv,fpthresq,fmois,ni,nj,nk,jcb8w,pbr,ptbr,ppp,ptpp,
 *            qvp,qallp,nsq8w,...)                       bruntv.f90:68-69
 *   s_defomten(...,jcb,jcb8u,jcb8v,jcb8w,...,up,vp,wp,t11,t22,t33,t12,t13,
 *              t23,t31,t32,...)                         defomten.f90:81-86
 *   s_defomssq(ni,nj,nk,t11,t22,t33,t12,t31,t32,ssq)     defomssq.f90:68
 *   s_eddyvis(...,dtb,ni,nj,nk,zph,jcb,rmf,rbr,tkep,ssq,nsq8w,rkh,rkv,priv)
 *                                                        eddyvis.f90:83-86
 * moist selects fmois 'moist' (else 'dry'); cfg->cphopt != 0 takes the
 * cloudy N^2 above cfg->thresq.  Arrays real(0:ni+1,0:nj+1,1:nk); the
 * outputs are written where the reference writes them and are otherwise
 * left as passed.  Periodic (-1) or open (7) sides.  The tensor, ssq and
 * dry / cloud-free N^2 are the FP32 reference's bits; rkh, rkv, priv and
 * cloudy N^2 (binary32 exp/log of their own, not libm's) are within 1e-5. */
int cress_turbcoef(const cress_config *cfg, float dtb, int moist, const float *jcb, const float *jcb8u,
                   const float *jcb8v, const float *jcb8w, const float *pbr, const float *ptbr, const float *rbr,
                   const float *ppp, const float *ptpp, const float *qvp, const float *qallp, const float *up,
                   const float *vp, const float *wp, const float *tkep, float *t11, float *t22, float *t33,
                   float *t12, float *t13, float *t23, float *t31, float *t32, float *ssq, float *nsq8w, float *rkh,
                   float *rkv, float *priv);

/* s_turbdrv(fmois,dtb,ni,nj,nk,zph,j31,j32,jcb,jcb8u,jcb8v,jcb8w,mf,mf8u,
 *           mf8v,rmf,rmf8u,rmf8v,pbr,ptbr,rbr,up,vp,wp,ppp,ptpp,qvp,tkep,qallp,
 *           ufrc,vfrc,wfrc,rkh,rkv,priv,ssq,nsq8w,t13,t23,t33,t11,t22,t12,t31,
 *           t32,tmp1)  SrcOrig/turbdrv.f90:80-85, for the options of
 * cress_turbcoef, then s_strsten (sfcopt=0) and s_turbuvw (trnopt=0,
 * mfcopt=0): ufrc, vfrc, wfrc are overwritten with the stress divergence
 * over their interiors; rkh, rkv, priv, ssq, nsq8w, t13, t23, t33 are
 * returned as the reference returns them (t11, t22, t12, t31, t32 and tmp1
 * are the reference's work arrays and are not taken).  The forcings are
 * within 1e-5 of the FP32 reference (they carry the eddy viscosity). */
int cress_turbdrv(const cress_config *cfg, float dtb, int moist, const float *jcb, const float *jcb8u,
                  const float *jcb8v, const float *jcb8w, const float *pbr, const float *ptbr, const float *rbr,
                  const float *up, const float *vp, const float *wp, const float *ppp, const float *ptpp,
                  const float *qvp, const float *tkep, const float *qallp, float *ufrc, float *vfrc, float *wfrc,
                  float *rkh, float *rkv, float *priv, float *ssq, float *nsq8w, float *t13, float *t23, float *t33);

/* s_turbflx(fptrnopt,fpsfcopt,fpdxiv,fpdyiv,fpdziv,ni,nj,nk,j31,j32,jcb,s,
 *           sfrc,rkh8u,rkh8v,rkv8w,h1,h2,h3,j31s,j32s,jcbs)
 * SrcOrig/turbflx.f90:72-74, trnopt=0 and sfcopt=0 (j31, j32, sfrc and the
 * work arrays are not read): the turbulent fluxes h1, h2, h3 of the scalar
 * s at u, v, w points, written where the reference writes them; bitwise. */
int cress_turbflx(const cress_config *cfg, const float *jcb, const float *s, const float *rkh8u,
                  const float *rkh8v, const float *rkv8w, float *h1, float *h2, float *h3);

/* s_turbs(fptrnopt,fpmpopt,fpmfcopt,fpdxiv,fpdyiv,fpdziv,ni,nj,nk,j31,j32,
 *         jcb8u,jcb8v,mf,rmf,rmf8u,rmf8v,h1,h2,h3,sfrc,tmp1,tmp2,tmp3)
 * SrcOrig/turbs.f90:72-75, trnopt=0 and mfcopt=0 (j31, j32 and the map
 * factors are not read): sfrc += the divergence of the fluxes h1, h2, h3
 * over i 2..ni-2, j 2..nj-2, k 2..nk-2; bitwise. */
int cress_turbs(const cress_config *cfg, const float *jcb8u, const float *jcb8v, const float *h1, const float *h2,
                const float *h3, float *sfrc);

/* s_kh8uv(fpmpopt,fpmfcopt,ni,nj,nk,rmf,rkh,rkh8u,rkh8v)
 * SrcOrig/kh8uv.f90:60, mfcopt=0 (rmf is not read): rkh8u(i) =
 * rkh(i-1)+rkh(i) over i 2..ni-1, j 1..nj-1, rkh8v likewise in j, k 1..nk-1;
 * bitwise. */
int cress_kh8uv(const cress_config *cfg, const float *rkh, float *rkh8u, float *rkh8v);

/* s_forcetke(fplspvar,fpvspvar,fplspopt,fpvspopt,fpsmtopt,ksp0,ni,nj,nk,
 *            j31,j32,jcb,jcb8u,jcb8v,jcb8w,mf,rmf,rmf8u,rmf8v,rbr,rst,rstxu,
 *            rstxv,rstxwc,tke,tkep,rbcxy,rbct,priv,nsq8w,ssq,rkh,rkv,tkefrc,
 *            h3,rkv8s,tmp1,tmp2,tmp3)  SrcOrig/forcetke.f90:88-94, for the
 * options of cress_turbdrv (tubopt=2, isoopt=1, trnopt=0, mfcopt=0), the
 * advection of cfg->advopt, no smoothing (cfg->smtopt % 10 == 0, else
 * CRESS_EUNSUPPORTED) and no lateral sponge.  The vertical sponge s_vsps0
 * acts when cfg->vspopt >= 1 and cfg->vspvar[9] == 'o', from level
 * cfg->ksp0[1]-1, with rbct(1:ni,1:nj,1:nk,1:2) (may be NULL otherwise).
 * tkefrc gets the TKE forcing (advection of tke, shear, dissipation,
 * diffusion of tkep, sponge, buoyancy); rkh, rkv come in as s_turbdrv
 * leaves them and go out as the scalar diffusivities the reference leaves
 * in them (rkv at w points), rkv8s with them.  ssq is read only (the
 * reference also uses it as a work array); j31, j32, the map factors, rbr,
 * rbcxy, h3 and tmp1..tmp3 are not taken.  rkh, rkv, rkv8s and all but the
 * dissipation term are the FP32 reference's bits; the dissipation's length
 * scale (a binary32 exp/log of its own) puts tkefrc within 1e-5. */
int cress_forcetke(const cress_config *cfg, const float *jcb, const float *jcb8u, const float *jcb8v,
                   const float *jcb8w, const float *rst, const float *rstxu, const float *rstxv, const float *rstxwc,
                   const float *tke, const float *tkep, const float *rbct, const float *priv, const float *nsq8w,
                   const float *ssq, float *rkh, float *rkv, float *tkefrc, float *rkv8s);

/* s_advp(fpadvopt,fpmpopt,fpmfcopt,fpdiaopt,fpiwest,...,fpdziv,ni,nj,nk,
 *        mf8u,mf8v,jcb8u,jcb8v,jcb8w,u,v,wc,pp,pfrc,jcbxu,jcbxv,jcbxwc,
 *        hadv,vadv,tmp1,tmp2,tmp3)  SrcOrig/advp.f90:75-79, mfcopt=0 and
 * diaopt=0 (the map factors are not read): writes the advection of p'
 * into pfrc over the interior. */
int cress_advp(const cress_config *cfg, const float *jcb8u, const float *jcb8v, const float *jcb8w,
               const float *u, const float *v, const float *wc, const float *pp, float *pfrc);

/* s_bc8u(fpwbc,fpebc,ni,nj,kmax,var8u)  SrcOrig/bc8u.f90:72,
 * s_bc8v(fpsbc,fpnbc,ni,nj,kmax,var8v)  SrcOrig/bc8v.f90:72,
 * s_bc8w(fpbbc,fptbc,ni,nj,nk,var8w)    SrcOrig/bc8w.f90:70:
 * the staggered-point fills at the outer boundaries of one rank, in place
 * on real(0:ni+1,0:nj+1,1:kmax) (1:nk for bc8w). */
int cress_bc8u(const cress_config *cfg, int kmax, float *var8u);
int cress_bc8v(const cress_config *cfg, int kmax, float *var8v);
int cress_bc8w(const cress_config *cfg, float *var8w);

/* s_bcycle(fpwbc,...,iwsnd,iwrcv,iesnd,iercv,jssnd,jsrcv,jnsnd,jnrcv,ni,nj,kmax,var)
 * SrcOrig/bcycle.f90:72-74: the one-rank periodic halo fill. */
int cress_bcycle(const cress_config *cfg, int iwsnd, int iwrcv, int iesnd, int iercv,
                 int jssnd, int jsrcv, int jnsnd, int jnrcv, int kmax, float *var);

#ifdef __cplusplus
}
#endif
#endif
