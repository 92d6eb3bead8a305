!-----------------------------------------------------------------------
! m_cress_hip -- ISO_C_BINDING interface to the C ABI of the MI355X large
! step (include/cress_hip.h).  A Fortran 90 host uses this module to drive
! the device-resident context (tier 1) or to call the per-routine shims
! (tier 2) with the reference's explicit-shape arrays.
!
! The derived type cress_config is interoperable with the C struct of the
! same name: same member order, C int / double / char kinds.
!-----------------------------------------------------------------------
module m_cress_hip

  use iso_c_binding

  implicit none

  private

  integer(c_int), parameter, public :: cress_ok = 0
  integer(c_int), parameter, public :: cress_einval = -1
  integer(c_int), parameter, public :: cress_enodev = -2
  integer(c_int), parameter, public :: cress_ehip = -3
  integer(c_int), parameter, public :: cress_enomem = -4
  integer(c_int), parameter, public :: cress_erange = -5
  integer(c_int), parameter, public :: cress_eunsupported = -6
  integer(c_int), parameter, public :: cress_etoobig = -7

! sizes of the result arrays of cress_diag, cress_autotune /
! cress_set_launch / cress_get_launch and cress_base_columns
  integer(c_int), parameter, public :: cress_ndiag = 13
  integer(c_int), parameter, public :: cress_nlaunch = 9
  integer(c_int), parameter, public :: cress_nsetupcol = 16

! field ids (base state: level 0 only; prognostics: 0 past, 1 present,
! 2 future)
  integer(c_int), parameter, public :: fid_jcb = 0, fid_jcb8u = 1,      &
     &  fid_jcb8v = 2, fid_jcb8w = 3, fid_ubr = 4, fid_vbr = 5,         &
     &  fid_pbr = 6, fid_ptbr = 7, fid_qvbr = 8, fid_rbr = 9,           &
     &  fid_rst = 10, fid_rst8u = 11, fid_rst8v = 12, fid_rst8w = 13,   &
     &  fid_rcsq = 14, fid_nbase = 15
  integer(c_int), parameter, public :: fid_u = 32, fid_v = 33,          &
     &  fid_w = 34, fid_pp = 35, fid_ptp = 36, fid_qv = 37, fid_qc = 38,&
     &  fid_qr = 39, fid_qi = 40, fid_qs = 41, fid_qg = 42, fid_nci = 43
  integer(c_int), parameter, public :: fid_fc = 64, fid_rbct = 65,      &
     &  fid_prwtr = 66, fid_price = 67

  type, bind(c), public :: cress_config
    integer(c_int) :: ni, nj, nk
    integer(c_int) :: wbc, ebc, sbc, nbc
    integer(c_int) :: bbc, tbc
    integer(c_int) :: advopt, impopt, gwmopt, buyopt, divopt
    integer(c_int) :: cphopt, haiopt, vspopt, coropt, sthopt
    integer(c_int) :: ksp0(2)
    character(kind=c_char) :: vspvar(10)
    integer(c_int) :: nsstp
    real(c_double) :: dx, dy, dz
    real(c_double) :: dtbig, dtsml
    real(c_double) :: weicoe, filcoe, thresq
    real(c_double) :: gwave
    integer(c_int) :: smtopt
    real(c_double) :: smhcoe, smvcoe
  end type cress_config

  public :: cress_error_string, cress_device_count
  public :: cress_ctx_create, cress_ctx_destroy, cress_field_size, cress_field_pitch
  public :: cress_upload, cress_download, cress_large_step, cress_sync
  public :: cress_set_timing, cress_stage_times
  public :: cress_gaussel, cress_advs, cress_bcycle, cress_advuvw, cress_advp
  public :: cress_hevi, cress_steps, cress_cloudphy, cress_turbcoef
  public :: cress_turbdrv, cress_turbflx, cress_turbs
  public :: cress_bc8u, cress_bc8v, cress_bc8w
  public :: cress_kh8uv, cress_gaussel_noalloc, cress_forcetke
  public :: cress_stepstke
  public :: cress_set_base, cress_base_columns, cress_diag
  public :: cress_autotune, cress_set_launch, cress_get_launch
  public :: cress_ens_create, cress_ens_step, cress_ens_sync, cress_ens_destroy
  public :: cress_mlp_create, cress_mlp_destroy, cress_mlp_init
  public :: cress_mlp_fit_norm, cress_mlp_train, cress_ctx_set_mlp
  public :: cress_ctx_mlp_cells
  public :: cress_errmsg, base_cfg

! a 1-D sounding for cress_set_base: z, u, v, p, pt, qv point to
! real(c_double) arrays of nlev+1 levels (heights increasing)
  type, bind(c), public :: cress_sounding
    integer(c_int) :: nlev
    type(c_ptr) :: z, u, v, p, pt, qv
    real(c_double) :: zsfc, dzmin, layer1, layer2, botbar, vspbar, fc
  end type cress_sounding

  interface

    function cress_error_string(code) bind(c, name='cress_error_string')
      import :: c_int, c_ptr
      integer(c_int), value :: code
      type(c_ptr) :: cress_error_string
    end function cress_error_string

    function cress_device_count(n) bind(c, name='cress_device_count')
      import :: c_int
      integer(c_int), intent(out) :: n
      integer(c_int) :: cress_device_count
    end function cress_device_count

    function cress_ctx_create(cfg, ctx) bind(c, name='cress_ctx_create')
      import :: c_int, c_ptr, cress_config
      type(cress_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: cress_ctx_create
    end function cress_ctx_create

    function cress_ctx_destroy(ctx) bind(c, name='cress_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: cress_ctx_destroy
    end function cress_ctx_destroy

    function cress_field_size(ctx, field, nfloat)                       &
     &  bind(c, name='cress_field_size')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      integer(c_size_t), intent(out) :: nfloat
      integer(c_int) :: cress_field_size
    end function cress_field_size

    function cress_field_pitch(ctx, row, plane)                         &
     &  bind(c, name='cress_field_pitch')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), intent(out) :: row, plane
      integer(c_int) :: cress_field_pitch
    end function cress_field_pitch

    function cress_upload(ctx, field, level, host)                      &
     &  bind(c, name='cress_upload')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, level
      real(c_float), intent(in) :: host(*)
      integer(c_int) :: cress_upload
    end function cress_upload

    function cress_download(ctx, field, level, host)                    &
     &  bind(c, name='cress_download')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, level
      real(c_float), intent(inout) :: host(*)
      integer(c_int) :: cress_download
    end function cress_download

    function cress_large_step(ctx, nsteps, first)                       &
     &  bind(c, name='cress_large_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, first
      integer(c_int) :: cress_large_step
    end function cress_large_step

    function cress_sync(ctx) bind(c, name='cress_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: cress_sync
    end function cress_sync

    function cress_set_timing(ctx, on) bind(c, name='cress_set_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      integer(c_int) :: cress_set_timing
    end function cress_set_timing

    function cress_stage_times(ctx, ms, nstage)                         &
     &  bind(c, name='cress_stage_times')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
      integer(c_int), value :: nstage
      integer(c_int) :: cress_stage_times
    end function cress_stage_times

! tier 2: s_gaussel (gaussel.f90:72-73), impopt=1
    function cress_gaussel(istr, iend, jstr, jend, kstr, kend,          &
     &  ni, nj, nk, rr, ss, tt, ff) bind(c, name='cress_gaussel')
      import :: c_int, c_float
      integer(c_int), value :: istr, iend, jstr, jend, kstr, kend
      integer(c_int), value :: ni, nj, nk
      real(c_float), intent(in) :: rr(*), tt(*)
      real(c_float), intent(inout) :: ss(*), ff(*)
      integer(c_int) :: cress_gaussel
    end function cress_gaussel

! tier 2: s_advs (advs.f90:75-77)
    function cress_advs(cfg, rstxu, rstxv, rstxwc, s, sfrc)             &
     &  bind(c, name='cress_advs')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: rstxu(*), rstxv(*), rstxwc(*), s(*)
      real(c_float), intent(inout) :: sfrc(*)
      integer(c_int) :: cress_advs
    end function cress_advs

! tier 2: s_advuvw (advuvw.f90:75-78)
    function cress_advuvw(cfg, rstxu, rstxv, rstxwc, u, v, w, ufrc, vfrc, &
     &  wfrc) bind(c, name='cress_advuvw')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: rstxu(*), rstxv(*), rstxwc(*)
      real(c_float), intent(in) :: u(*), v(*), w(*)
      real(c_float), intent(inout) :: ufrc(*), vfrc(*), wfrc(*)
      integer(c_int) :: cress_advuvw
    end function cress_advuvw

! tier 2: s_hevi (hevi.f90:88-95), one substep, periodic sides
    function cress_hevi(cfg, dts, jcb, jcb8u, jcb8v, jcb8w, rbr, rst,    &
     &  rst8u, rst8v, rst8w, rcsq, ufrc, vfrc, wfrc, pfrc, uf, vf, wf, wc, &
     &  ppf) bind(c, name='cress_hevi')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), value :: dts
      real(c_float), intent(in) :: jcb(*), jcb8u(*), jcb8v(*), jcb8w(*)
      real(c_float), intent(in) :: rbr(*), rst(*), rst8u(*), rst8v(*)
      real(c_float), intent(in) :: rst8w(*), rcsq(*)
      real(c_float), intent(in) :: ufrc(*), vfrc(*), wfrc(*), pfrc(*)
      real(c_float), intent(inout) :: uf(*), vf(*), wf(*), wc(*), ppf(*)
      integer(c_int) :: cress_hevi
    end function cress_hevi

! tier 2: s_steps (steps.f90:114-127), periodic sides
    function cress_steps(cfg, dtb, rst, ptpp, ptfrc, qvp, qvfrc, qwtrp,  &
     &  qicep, nicep, qwtrf, qicef, nicef, ptpf, qvf, dtdrst)              &
     &  bind(c, name='cress_steps')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), value :: dtb
      real(c_float), intent(in) :: rst(*), ptpp(*), ptfrc(*), qvp(*)
      real(c_float), intent(in) :: qvfrc(*), qwtrp(*), qicep(*), nicep(*)
      real(c_float), intent(inout) :: qwtrf(*), qicef(*), nicef(*)
      real(c_float), intent(inout) :: ptpf(*), qvf(*), dtdrst(*)
      integer(c_int) :: cress_steps
    end function cress_steps

! tier 2: s_cloudphy (cloudphy.f90:84-91), cphopt=2, haiopt=0
    function cress_cloudphy(cfg, dtb, jcb, pbr, ptbr, rbr, rst, ppp,     &
     &  ptpp, qvp, qwtrp, qicep, nicep, ptpf, qvf, qwtrf, qicef, nicef,  &
     &  prwtr, price) bind(c, name='cress_cloudphy')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), value :: dtb
      real(c_float), intent(in) :: jcb(*), pbr(*), ptbr(*), rbr(*)
      real(c_float), intent(in) :: rst(*), ppp(*), ptpp(*), qvp(*)
      real(c_float), intent(in) :: qwtrp(*), qicep(*), nicep(*)
      real(c_float), intent(inout) :: ptpf(*), qvf(*), qwtrf(*)
      real(c_float), intent(inout) :: qicef(*), nicef(*)
      real(c_float), intent(inout) :: prwtr(*), price(*)
      integer(c_int) :: cress_cloudphy
    end function cress_cloudphy

! tier 2: the first four calls of s_turbdrv (turbdrv.f90:245-276), tubopt=2
    function cress_turbcoef(cfg, dtb, moist, jcb, jcb8u, jcb8v, jcb8w,   &
     &  pbr, ptbr, rbr, ppp, ptpp, qvp, qallp, up, vp, wp, tkep, t11,     &
     &  t22, t33, t12, t13, t23, t31, t32, ssq, nsq8w, rkh, rkv, priv)    &
     &  bind(c, name='cress_turbcoef')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), value :: dtb
      integer(c_int), value :: moist
      real(c_float), intent(in) :: jcb(*), jcb8u(*), jcb8v(*), jcb8w(*)
      real(c_float), intent(in) :: pbr(*), ptbr(*), rbr(*), ppp(*), ptpp(*)
      real(c_float), intent(in) :: qvp(*), qallp(*), up(*), vp(*), wp(*)
      real(c_float), intent(in) :: tkep(*)
      real(c_float), intent(inout) :: t11(*), t22(*), t33(*), t12(*)
      real(c_float), intent(inout) :: t13(*), t23(*), t31(*), t32(*)
      real(c_float), intent(inout) :: ssq(*), nsq8w(*), rkh(*), rkv(*)
      real(c_float), intent(inout) :: priv(*)
      integer(c_int) :: cress_turbcoef
    end function cress_turbcoef

! tier 2: s_turbdrv (turbdrv.f90:80-85), tubopt=2
    function cress_turbdrv(cfg, dtb, moist, jcb, jcb8u, jcb8v, jcb8w,    &
     &  pbr, ptbr, rbr, up, vp, wp, ppp, ptpp, qvp, tkep, qallp, ufrc,    &
     &  vfrc, wfrc, rkh, rkv, priv, ssq, nsq8w, t13, t23, t33)            &
     &  bind(c, name='cress_turbdrv')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), value :: dtb
      integer(c_int), value :: moist
      real(c_float), intent(in) :: jcb(*), jcb8u(*), jcb8v(*), jcb8w(*)
      real(c_float), intent(in) :: pbr(*), ptbr(*), rbr(*), up(*), vp(*)
      real(c_float), intent(in) :: wp(*), ppp(*), ptpp(*), qvp(*), tkep(*)
      real(c_float), intent(in) :: qallp(*)
      real(c_float), intent(inout) :: ufrc(*), vfrc(*), wfrc(*), rkh(*)
      real(c_float), intent(inout) :: rkv(*), priv(*), ssq(*), nsq8w(*)
      real(c_float), intent(inout) :: t13(*), t23(*), t33(*)
      integer(c_int) :: cress_turbdrv
    end function cress_turbdrv

! tier 2: s_turbflx (turbflx.f90:72-74), trnopt=0, sfcopt=0
    function cress_turbflx(cfg, jcb, s, rkh8u, rkh8v, rkv8w, h1, h2, h3) &
     &  bind(c, name='cress_turbflx')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: jcb(*), s(*), rkh8u(*), rkh8v(*), rkv8w(*)
      real(c_float), intent(inout) :: h1(*), h2(*), h3(*)
      integer(c_int) :: cress_turbflx
    end function cress_turbflx

! tier 2: s_turbs (turbs.f90:72-75), trnopt=0, mfcopt=0
    function cress_turbs(cfg, jcb8u, jcb8v, h1, h2, h3, sfrc)            &
     &  bind(c, name='cress_turbs')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: jcb8u(*), jcb8v(*), h1(*), h2(*), h3(*)
      real(c_float), intent(inout) :: sfrc(*)
      integer(c_int) :: cress_turbs
    end function cress_turbs

! tier 2: s_advp (advp.f90:75-79), mfcopt=0, diaopt=0
    function cress_advp(cfg, jcb8u, jcb8v, jcb8w, u, v, wc, pp, pfrc)    &
     &  bind(c, name='cress_advp')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: jcb8u(*), jcb8v(*), jcb8w(*)
      real(c_float), intent(in) :: u(*), v(*), wc(*), pp(*)
      real(c_float), intent(inout) :: pfrc(*)
      integer(c_int) :: cress_advp
    end function cress_advp

! tier 2: s_bc8u, s_bc8v (bc8u.f90:72, bc8v.f90:72), s_bc8w (bc8w.f90:70)
    function cress_bc8u(cfg, kmax, var) bind(c, name='cress_bc8u')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      integer(c_int), value :: kmax
      real(c_float), intent(inout) :: var(*)
      integer(c_int) :: cress_bc8u
    end function cress_bc8u

    function cress_bc8v(cfg, kmax, var) bind(c, name='cress_bc8v')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      integer(c_int), value :: kmax
      real(c_float), intent(inout) :: var(*)
      integer(c_int) :: cress_bc8v
    end function cress_bc8v

    function cress_bc8w(cfg, var) bind(c, name='cress_bc8w')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(inout) :: var(*)
      integer(c_int) :: cress_bc8w
    end function cress_bc8w

! tier 2: s_bcycle (bcycle.f90:72-74)
    function cress_bcycle(cfg, iwsnd, iwrcv, iesnd, iercv,              &
     &  jssnd, jsrcv, jnsnd, jnrcv, kmax, var) bind(c, name='cress_bcycle')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      integer(c_int), value :: iwsnd, iwrcv, iesnd, iercv
      integer(c_int), value :: jssnd, jsrcv, jnsnd, jnrcv, kmax
      real(c_float), intent(inout) :: var(*)
      integer(c_int) :: cress_bcycle
    end function cress_bcycle

! tier 2: s_kh8uv (kh8uv.f90:60), mfcopt=0
    function cress_kh8uv(cfg, rkh, rkh8u, rkh8v) bind(c, name='cress_kh8uv')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: rkh(*)
      real(c_float), intent(inout) :: rkh8u(*), rkh8v(*)
      integer(c_int) :: cress_kh8uv
    end function cress_kh8uv

! tier 2: s_forcetke (forcetke.f90:88-94), tubopt=2, isoopt=1, trnopt=0,
! mfcopt=0, smtopt=0, no lateral sponge
    function cress_forcetke(cfg, jcb, jcb8u, jcb8v, jcb8w, rst, rstxu,  &
     &  rstxv, rstxwc, tke, tkep, rbct, priv, nsq8w, ssq, rkh, rkv,      &
     &  tkefrc, rkv8s) bind(c, name='cress_forcetke')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), intent(in) :: jcb(*), jcb8u(*), jcb8v(*), jcb8w(*)
      real(c_float), intent(in) :: rst(*), rstxu(*), rstxv(*), rstxwc(*)
      real(c_float), intent(in) :: tke(*), tkep(*), rbct(*), priv(*)
      real(c_float), intent(in) :: nsq8w(*), ssq(*)
      real(c_float), intent(inout) :: rkh(*), rkv(*), tkefrc(*), rkv8s(*)
      integer(c_int) :: cress_forcetke
    end function cress_forcetke

! tier 2: the TKE lines of s_steps (steps.f90:997-1012), periodic sides
    function cress_stepstke(cfg, dtb, rst, tkep, tkef)                  &
     &  bind(c, name='cress_stepstke')
      import :: c_int, c_float, cress_config
      type(cress_config), intent(in) :: cfg
      real(c_float), value :: dtb
      real(c_float), intent(in) :: rst(*), tkep(*)
      real(c_float), intent(inout) :: tkef(*)
      integer(c_int) :: cress_stepstke
    end function cress_stepstke

! cress_gaussel without device allocation (a fixed device array stages the
! four arrays; larger boxes return cress_etoobig)
    function cress_gaussel_noalloc(istr, iend, jstr, jend, kstr, kend,  &
     &  ni, nj, nk, rr, ss, tt, ff) bind(c, name='cress_gaussel_noalloc')
      import :: c_int, c_float
      integer(c_int), value :: istr, iend, jstr, jend, kstr, kend
      integer(c_int), value :: ni, nj, nk
      real(c_float), intent(in) :: rr(*), tt(*)
      real(c_float), intent(inout) :: ss(*), ff(*)
      integer(c_int) :: cress_gaussel_noalloc
    end function cress_gaussel_noalloc

! base state from a sounding in binary64 on the GPU (stretch, Jacobian,
! intrpsnd, setbase, baserho, var8uvw, sndwave, vspdmp)
    function cress_set_base(ctx, snd) bind(c, name='cress_set_base')
      import :: c_int, c_ptr, cress_sounding
      type(c_ptr), value :: ctx
      type(cress_sounding), intent(in) :: snd
      integer(c_int) :: cress_set_base
    end function cress_set_base

    function cress_base_columns(ctx, out, ncol)                        &
     &  bind(c, name='cress_base_columns')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: ncol
      integer(c_int) :: cress_base_columns
    end function cress_base_columns

! binary64 budgets of a level (cress_ndiag values)
    function cress_diag(ctx, level, out) bind(c, name='cress_diag')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: cress_diag
    end function cress_diag

! launch shapes (cress_nlaunch integers)
    function cress_autotune(ctx, shape, n) bind(c, name='cress_autotune')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: shape(*)
      integer(c_int), value :: n
      integer(c_int) :: cress_autotune
    end function cress_autotune

    function cress_set_launch(ctx, shape, n)                           &
     &  bind(c, name='cress_set_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: shape(*)
      integer(c_int), value :: n
      integer(c_int) :: cress_set_launch
    end function cress_set_launch

    function cress_get_launch(ctx, shape, n)                           &
     &  bind(c, name='cress_get_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: shape(*)
      integer(c_int), value :: n
      integer(c_int) :: cress_get_launch
    end function cress_get_launch

! ensembles: members(1:n) are contexts
    function cress_ens_create(members, n, e)                           &
     &  bind(c, name='cress_ens_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: members(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: e
      integer(c_int) :: cress_ens_create
    end function cress_ens_create

    function cress_ens_step(e, nsteps, first) bind(c, name='cress_ens_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps, first
      integer(c_int) :: cress_ens_step
    end function cress_ens_step

    function cress_ens_sync(e) bind(c, name='cress_ens_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int) :: cress_ens_sync
    end function cress_ens_sync

    function cress_ens_destroy(e) bind(c, name='cress_ens_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int) :: cress_ens_destroy
    end function cress_ens_destroy

! the learned microphysics
    function cress_mlp_create(capacity, max_batch, m)                  &
     &  bind(c, name='cress_mlp_create')
      import :: c_int, c_long, c_ptr
      integer(c_long), value :: capacity
      integer(c_int), value :: max_batch
      type(c_ptr), intent(out) :: m
      integer(c_int) :: cress_mlp_create
    end function cress_mlp_create

    function cress_mlp_destroy(m) bind(c, name='cress_mlp_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int) :: cress_mlp_destroy
    end function cress_mlp_destroy

    function cress_mlp_init(m, seed) bind(c, name='cress_mlp_init')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: m
      integer(c_long_long), value :: seed
      integer(c_int) :: cress_mlp_init
    end function cress_mlp_init

    function cress_mlp_fit_norm(m, margin) bind(c, name='cress_mlp_fit_norm')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), value :: margin
      integer(c_int) :: cress_mlp_fit_norm
    end function cress_mlp_fit_norm

    function cress_mlp_train(m, steps, batch, lr, seed, losses)        &
     &  bind(c, name='cress_mlp_train')
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: m
      integer(c_int), value :: steps, batch
      real(c_float), value :: lr
      integer(c_int), value :: seed
      real(c_double), intent(out) :: losses(*)
      integer(c_int) :: cress_mlp_train
    end function cress_mlp_train

    function cress_ctx_set_mlp(ctx, m, mode) bind(c, name='cress_ctx_set_mlp')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, m
      integer(c_int), value :: mode
      integer(c_int) :: cress_ctx_set_mlp
    end function cress_ctx_set_mlp

    function cress_ctx_mlp_cells(ctx, cells) bind(c, name='cress_ctx_mlp_cells')
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: ctx
      integer(c_long), intent(out) :: cells
      integer(c_int) :: cress_ctx_mlp_cells
    end function cress_ctx_mlp_cells

  end interface

contains

! A configuration for the per-routine shims: the given dimensions,
! advection order and grid spacings, periodic sides, rigid top and bottom,
! the plain-path defaults elsewhere (the shims read only what their
! routine uses).
  subroutine base_cfg(cfg, ni, nj, nk, advopt, dx, dy, dz)
    type(cress_config), intent(out) :: cfg
    integer, intent(in) :: ni, nj, nk, advopt
    real, intent(in) :: dx, dy, dz
    cfg%ni = ni
    cfg%nj = nj
    cfg%nk = nk
    cfg%wbc = -1
    cfg%ebc = -1
    cfg%sbc = -1
    cfg%nbc = -1
    cfg%bbc = 2
    cfg%tbc = 2
    cfg%advopt = advopt
    cfg%impopt = 1
    cfg%gwmopt = 0
    cfg%buyopt = 1
    cfg%divopt = 1
    cfg%cphopt = 0
    cfg%haiopt = 0
    cfg%vspopt = 0
    cfg%coropt = 0
    cfg%sthopt = 0
    cfg%ksp0(1) = 2
    cfg%ksp0(2) = 2
    cfg%vspvar(1:10) = 'x'
    cfg%nsstp = 1
    cfg%dx = dx
    cfg%dy = dy
    cfg%dz = dz
    cfg%dtbig = 1.d0
    cfg%dtsml = 1.d0
    cfg%weicoe = .6d0
    cfg%filcoe = .1d0
    cfg%thresq = 1.d-8
    cfg%gwave = 0.d0
    cfg%smtopt = 0
    cfg%smhcoe = 0.d0
    cfg%smvcoe = 0.d0
  end subroutine base_cfg

! The message of a return code as a Fortran string.
  function cress_errmsg(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=64) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: n

    msg = ' '
    p = cress_error_string(code)
    if(.not.c_associated(p)) return
    call c_f_pointer(p, cs, [64])
    do n = 1, 64
      if(cs(n) == c_null_char) exit
      msg(n:n) = cs(n)
    end do
  end function cress_errmsg

end module m_cress_hip
