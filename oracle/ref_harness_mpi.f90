! Test-infrastructure harness, part 2: bind(C) entry points into the
! reference CReSS routines whose module closure reaches the halo-exchange
! layer (m_defmpi): the whole large step s_intgdrv, the acoustic substep
! s_hevi, s_stepuv, s_stepwi, s_steppi, s_steps, s_cloudphy, s_fallblk,
! s_phy2cnt and s_shift2nd.  They are called unchanged at one rank
! (numpe=1), where every halo exchange is a no-op (shiftsx.f90:139,
! putbufsx.f90:137) and the MPI reductions (chkfall.f90:88,
! chkerr_multi.f90:95) run over one process.  Linked only into
! oracle/_ref/libcressrefmpi32.so, with MPICH from the image.
!
! Arrays follow the reference's explicit-shape real(0:ni+1,0:nj+1,1:nk)
! layout.  The caller passes three blocks:
!   base(:,:,:,1:15)  jcb,jcb8u,jcb8v,jcb8w,ubr,vbr,pbr,ptbr,qvbr,rbr,rst,
!                     rst8u,rst8v,rst8w,rcsq
!   lev(:,:,:,1:16,1:3)  per level (1 past, 2 present, 3 future):
!                     u,v,w,pp,ptp,qv,qc,qr,qi,qs,qg,ncc,ncr,nci,ncs,ncg
!                     so that qwtr=(qc,qr), qice=(qi,qs,qg), nwtr=(ncc,ncr),
!                     nice=(nci,ncs,ncg) are contiguous as the reference
!                     stores them (allocslv.f90);
!   fc(0:ni+1,0:nj+1,1:2), rbct(1:ni,1:nj,1:nk,1:2),
!   prwtr(0:ni+1,0:nj+1,1:2,1:2), price(0:ni+1,0:nj+1,1:2,1:3).
! Everything the plain path leaves unused (terrain slopes, map factors of 1,
! nudging and external-boundary data, aerosol, tracer, TKE, soil) is held
! here, zeroed, as are the open-boundary phase speeds, which persist from
! one call to the next as they do in the reference's allocslv arrays.

      module cress_ref_mpi

      use iso_c_binding
      use m_comindx
      use m_comname
      use m_commpi
      use m_inimpi
      use m_allocmph
      use m_setref
      use m_intgdrv
      use m_hevi
      use m_stepuv
      use m_stepwi
      use m_steppi
      use m_steps
      use m_cloudphy
      use m_fallblk
      use m_phy2cnt
      use m_shift2nd
      use m_stretch
      use m_intrpsnd
      use m_setbase
      use m_baserho
      use m_var8uvw
      use m_sndwave
      use m_vspdmp

      implicit none

      integer, parameter :: i8 = selected_int_kind(18)

      integer, save :: mi = -1, mj = -1, mk = -1
      integer, save :: nqw, nnw, nqi, nni, nqa(0:4)

      real, allocatable, save :: z3(:,:,:), zmf(:,:), zrmf(:,:,:)
      real, allocatable, save :: zero3(:,:,:,:), zero2(:,:,:)
      real, allocatable, save :: cpx(:,:,:,:), cpy(:,:,:,:)
      real, allocatable, save :: wk(:,:,:,:)
      integer, allocatable, save :: land(:,:)

      contains

! MPI (singleton at one rank) and the reference's tables of constants.
      subroutine ref_mpi_init() bind(C,name='ref_mpi_init')
      call s_inimpi('solver  ',6)
      call s_setref
      end subroutine ref_mpi_init

! Options ref_configure does not set, and the workspace for one grid.
! Call after ref_configure; the grid may not change within a process
! (allocmph allocates the reference's module arrays once).
      subroutine ref_mpi_setup(ni,nj,nk,sthopt,cphopt,gwave,smtopt,smh,   &
     &   smv) bind(C,name='ref_mpi_setup')
      integer(c_int), value :: ni,nj,nk,sthopt,cphopt,smtopt
      real, value :: gwave,smh,smv
      integer stat
      iname(idzeropt)=0
      iname(idoneopt)=1
      iname(idsthopt)=sthopt
      rname(idgwave)=gwave
      iname(idsmtopt)=smtopt
      rname(idsmhcoe)=smh
      rname(idsmvcoe)=smv
      iname(idsavmem)=0
      if(abs(cphopt).eq.0) then
        nqw=1
        nnw=1
        nqi=1
        nni=1
      else
        nqw=2
        nnw=2
        nqi=3
        nni=3
      end if
      nqa(0)=1
      nqa(1:4)=0
      mi=ni
      mj=nj
      mk=nk
      allocate(z3(0:ni+1,0:nj+1,1:nk),stat=stat)
      allocate(zmf(0:ni+1,0:nj+1),zrmf(0:ni+1,0:nj+1,1:4))
      allocate(zero3(0:ni+1,0:nj+1,1:nk,1:40),zero2(0:ni+1,0:nj+1,1:16))
      allocate(cpx(1:nj,1:nk,1:2,1:40),cpy(1:ni,1:nk,1:2,1:40))
      allocate(wk(0:ni+1,0:nj+1,1:nk,1:24))
      allocate(land(0:ni+1,0:nj+1))
      z3(:,:,:)=0.e0
      zmf(:,:)=1.e0
      zrmf(:,:,:)=1.e0
      zero3(:,:,:,:)=0.e0
      zero2(:,:,:)=0.e0
      cpx(:,:,:,:)=0.e0
      cpy(:,:,:,:)=0.e0
      wk(:,:,:,:)=0.e0
      land(:,:)=0
      if(abs(cphopt).ge.2) then
        call s_allocmph(idsavmem,idcphopt,idhaiopt,idaslopt,ni,nj,nk,nqw)
      end if
      end subroutine ref_mpi_setup

! Fill the work arrays the harness owns (the reference's tmp1..tmp7,
! forcings, wc, qall, qallp) with one value.
      subroutine ref_mpi_fill_work(v) bind(C,name='ref_mpi_fill_work')
      real, value :: v
      wk(:,:,:,:)=v
      end subroutine ref_mpi_fill_work

! Copy work array number k (1..24) out: after ref_intgdrv, 5..9 hold the
! step's ufrc, vfrc, wfrc, ptfrc, qvfrc, 3 pfrc and 4 wc.
      subroutine ref_mpi_get_work(k,out) bind(C,name='ref_mpi_get_work')
      integer(c_int), value :: k
      real out(0:mi+1,0:mj+1,1:mk)
      out(:,:,:)=wk(:,:,:,k)
      end subroutine ref_mpi_get_work

! Largest magnitude held in the harness's zero-filled arrays (which the
! plain path must leave untouched), for checks.
      real(c_float) function ref_mpi_dirty() bind(C,name='ref_mpi_dirty')
      ref_mpi_dirty=max(maxval(abs(z3)),maxval(abs(zero3)),              &
     &  maxval(abs(zero2)),maxval(abs(zmf-1.e0)),maxval(abs(zrmf-1.e0)))
      end function ref_mpi_dirty

! One large step, s_intgdrv (intgdrv.f90:820-1133), with the time
! bookkeeping s_timeint gives it (timeint.f90:160-222): ctime is the model
! time in ms (0 for the forward first step), dtb/dts the step's intervals.
      subroutine ref_intgdrv(ni,nj,nk,ksp0a,ksp0b,nsstp,ctime,dtb,dts,   &
     &   moist,base,fc,rbct,lev,prwtr,price) bind(C,name='ref_intgdrv')
      integer(c_int), value :: ni,nj,nk,ksp0a,ksp0b,nsstp,moist
      integer(c_int64_t), value :: ctime
      real, value :: dtb,dts
      real base(0:ni+1,0:nj+1,1:nk,1:15)
      real fc(0:ni+1,0:nj+1,1:2),rbct(1:ni,1:nj,1:nk,1:2)
      real lev(0:ni+1,0:nj+1,1:nk,1:16,1:3)
      real prwtr(0:ni+1,0:nj+1,1:2,1:2),price(0:ni+1,0:nj+1,1:2,1:3)
      character(len=5) fmois
      character(len=12) pdate
      integer ksp0(1:2),nclstp(0:3)
      integer(kind=i8) ct,ft
      real dtcl(1:3),ngrdmp(1:2),rtinc(1:2),area(0:4),dtsep,dtsoil
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_intgdrv: setup'
      if(moist.ne.0) then
        fmois='moist'
      else
        fmois='dry  '
      end if
      pdate='000000000000'
      ksp0(1)=ksp0a
      ksp0(2)=ksp0b
      nclstp(:)=0
      ct=int(ctime,i8)
      ft=ct+int(1000.e0*dtb,i8)
      dtcl(:)=dtb
      ngrdmp(:)=0.e0
      rtinc(:)=0.e0
      area(:)=0.e0
      dtsep=-.1e0
      dtsoil=-.1e0
      call s_intgdrv(idexbopt,idsfcopt,idadvopt,idcphopt,idaslopt,      &
     &  idtubopt,iddiaopt,fmois,pdate,ksp0,nsstp,1,nclstp,              &
     &  ct,ft,dtb,dts,dtsep,dtsoil,dtcl,0.e0,ngrdmp,0.e0,0.e0,rtinc,    &
     &  0.e0,area,ni,nj,nk,nqw,nnw,nqi,nni,nk,nqa,1,                    &
     &  z3,zmf,zmf,z3,z3,                                               &
     &  base(0,0,1,1),base(0,0,1,2),base(0,0,1,3),base(0,0,1,4),        &
     &  zmf,zmf,zmf,zrmf,zrmf,zrmf,fc,                                  &
     &  base(0,0,1,5),base(0,0,1,6),base(0,0,1,7),base(0,0,1,8),        &
     &  base(0,0,1,9),base(0,0,1,10),base(0,0,1,11),base(0,0,1,12),     &
     &  base(0,0,1,13),base(0,0,1,14),base(0,0,1,15),                   &
     &  zero2(1,1,1),zero2(1,1,1),zero2(1,1,1),rbct,                    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),                                  &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),                   &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  land,zero2(0,0,1),zero2(0,0,2),zero2(0,0,3),zero2(0,0,4),       &
     &  zero2(0,0,5),zero2(0,0,6),zero2(0,0,7),                         &
     &  lev(0,0,1,1,2),lev(0,0,1,1,1),lev(0,0,1,2,2),lev(0,0,1,2,1),    &
     &  lev(0,0,1,3,2),lev(0,0,1,3,1),lev(0,0,1,4,2),lev(0,0,1,4,1),    &
     &  lev(0,0,1,5,2),lev(0,0,1,5,1),lev(0,0,1,6,2),lev(0,0,1,6,1),    &
     &  lev(0,0,1,7,2),lev(0,0,1,7,1),lev(0,0,1,12,2),lev(0,0,1,12,1),  &
     &  lev(0,0,1,9,2),lev(0,0,1,9,1),lev(0,0,1,14,2),lev(0,0,1,14,1),  &
     &  zero3(0,0,1,11),zero3(0,0,1,14),zero3(0,0,1,17),zero3(0,0,1,20),&
     &  zero3(0,0,1,23),zero3(0,0,1,24),zero3(0,0,1,25),zero3(0,0,1,26),&
     &  zero3(0,0,1,27),zero3(0,0,1,28),                                &
     &  cpx(1,1,1,1),cpy(1,1,1,1),cpx(1,1,1,2),cpy(1,1,1,2),            &
     &  cpx(1,1,1,3),cpy(1,1,1,3),cpx(1,1,1,4),cpy(1,1,1,4),            &
     &  cpx(1,1,1,5),cpy(1,1,1,5),cpx(1,1,1,6),cpy(1,1,1,6),            &
     &  cpx(1,1,1,7),cpy(1,1,1,7),cpx(1,1,1,9),cpy(1,1,1,9),            &
     &  cpx(1,1,1,11),cpy(1,1,1,11),cpx(1,1,1,14),cpy(1,1,1,14),        &
     &  cpx(1,1,1,17),cpy(1,1,1,17),cpx(1,1,1,19),cpy(1,1,1,19),        &
     &  cpx(1,1,1,22),cpy(1,1,1,22),cpx(1,1,1,23),cpy(1,1,1,23),        &
     &  cpx(1,1,1,24),cpy(1,1,1,24),                                    &
     &  prwtr,price,wk(0,0,1,1),wk(0,0,1,2),wk(0,0,1,3),                &
     &  zero2(0,0,8),zero2(0,0,9),zero2(0,0,10),zero2(0,0,11),          &
     &  lev(0,0,1,1,3),lev(0,0,1,2,3),lev(0,0,1,3,3),wk(0,0,1,4),       &
     &  lev(0,0,1,4,3),lev(0,0,1,5,3),lev(0,0,1,6,3),lev(0,0,1,7,3),    &
     &  lev(0,0,1,12,3),lev(0,0,1,9,3),lev(0,0,1,14,3),                 &
     &  zero3(0,0,1,29),zero3(0,0,1,32),zero3(0,0,1,35),zero3(0,0,1,36),&
     &  zero3(0,0,1,37),zero3(0,0,1,38),                                &
     &  wk(0,0,1,5),wk(0,0,1,6),wk(0,0,1,7),wk(0,0,1,8),wk(0,0,1,9),    &
     &  wk(0,0,1,10),wk(0,0,1,11),wk(0,0,1,12),wk(0,0,1,13),            &
     &  wk(0,0,1,14),wk(0,0,1,15),wk(0,0,1,16))
      end subroutine ref_intgdrv

! One acoustic substep, s_hevi (hevi.f90:88-431), on the future level
! fut(:,:,:,1:6) = uf,vf,wf,wc,ppf,ptpf with the forcings
! frc(:,:,:,1:5) = ufrc,vfrc,wfrc,pfrc,ptfrc; tmp(:,:,:,1:6) work.
      subroutine ref_hevi(ni,nj,nk,isstp,dts,base,frc,fut,tmp)           &
     &   bind(C,name='ref_hevi')
      integer(c_int), value :: ni,nj,nk,isstp
      real, value :: dts
      real base(0:ni+1,0:nj+1,1:nk,1:15),frc(0:ni+1,0:nj+1,1:nk,1:5)
      real fut(0:ni+1,0:nj+1,1:nk,1:6),tmp(0:ni+1,0:nj+1,1:nk,1:6)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_hevi: setup'
      call s_hevi(idgwmopt,idbuyopt,isstp,dts,0.e0,ni,nj,nk,z3,z3,       &
     &  base(0,0,1,1),base(0,0,1,2),base(0,0,1,3),base(0,0,1,4),        &
     &  zmf,zmf,zmf,zrmf,zrmf,zrmf,                                     &
     &  base(0,0,1,5),base(0,0,1,6),base(0,0,1,8),base(0,0,1,10),       &
     &  base(0,0,1,11),base(0,0,1,12),base(0,0,1,13),base(0,0,1,14),    &
     &  base(0,0,1,15),frc(0,0,1,1),frc(0,0,1,2),frc(0,0,1,3),          &
     &  frc(0,0,1,4),frc(0,0,1,5),                                      &
     &  cpx(1,1,1,1),cpy(1,1,1,1),cpx(1,1,1,2),cpy(1,1,1,2),            &
     &  cpx(1,1,1,3),cpy(1,1,1,3),cpx(1,1,1,4),cpy(1,1,1,4),            &
     &  cpx(1,1,1,5),cpy(1,1,1,5),                                      &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),                                  &
     &  fut(0,0,1,1),fut(0,0,1,2),fut(0,0,1,3),fut(0,0,1,4),            &
     &  fut(0,0,1,5),fut(0,0,1,6),tmp(0,0,1,1),tmp(0,0,1,2),            &
     &  tmp(0,0,1,3),tmp(0,0,1,4),tmp(0,0,1,5),tmp(0,0,1,6))
      end subroutine ref_hevi

! s_stepuv (stepuv.f90:223-424): uf,vf += dts*(frc+pgf)/rst8 with the
! lateral and vertical fills.
      subroutine ref_stepuv(ni,nj,nk,isstp,dts,rst8u,rst8v,ufrc,vfrc,    &
     &   usml,vsml,uf,vf) bind(C,name='ref_stepuv')
      integer(c_int), value :: ni,nj,nk,isstp
      real, value :: dts
      real rst8u(*),rst8v(*),ufrc(*),vfrc(*),usml(*),vsml(*),uf(*),vf(*)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_stepuv: setup'
      call s_stepuv(idexbvar,idexbopt,isstp,dts,0.e0,ni,nj,nk,z3,z3,     &
     &  rst8u,rst8v,ufrc,vfrc,usml,vsml,cpx(1,1,1,1),cpy(1,1,1,1),      &
     &  cpx(1,1,1,2),cpy(1,1,1,2),zero3(0,0,1,1),zero3(0,0,1,1),        &
     &  zero3(0,0,1,1),zero3(0,0,1,1),uf,vf)
      end subroutine ref_stepuv

! s_stepwi (stepwi.f90:108-553): explicit w predictor, the implicit w-p
! tridiagonal (gaussel), the fills and wc.
      subroutine ref_stepwi(ni,nj,nk,isstp,dts,jcb,jcb8w,rbr,rst,rst8w,  &
     &   rcsq,uf,vf,fw,wf,wc,t1,t2,t3) bind(C,name='ref_stepwi')
      integer(c_int), value :: ni,nj,nk,isstp
      real, value :: dts
      real jcb(*),jcb8w(*),rbr(*),rst(*),rst8w(*),rcsq(*),uf(*),vf(*)
      real fw(*),wf(*),wc(*),t1(*),t2(*),t3(*)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_stepwi: setup'
      call s_stepwi(idexbvar,idexbopt,idimpopt,idbuyopt,iddziv,idweicoe, &
     &  isstp,dts,0.e0,ni,nj,nk,z3,z3,jcb,jcb8w,zmf,rbr,rst,rst8w,rcsq,  &
     &  uf,vf,cpx(1,1,1,3),cpy(1,1,1,3),zero3(0,0,1,1),zero3(0,0,1,1),  &
     &  fw,wf,wc,t1,t2,t3)
      end subroutine ref_stepwi

! s_steppi (steppi.f90:174-316): ppf += dts*fp/jcb with the fills.
      subroutine ref_steppi(ni,nj,nk,isstp,dts,jcb,fp,ppf)               &
     &   bind(C,name='ref_steppi')
      integer(c_int), value :: ni,nj,nk,isstp
      real, value :: dts
      real jcb(*),fp(*),ppf(*)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_steppi: setup'
      call s_steppi(idexbvar,idexbopt,isstp,dts,0.e0,ni,nj,nk,jcb,fp,    &
     &  cpx(1,1,1,4),cpy(1,1,1,4),zero3(0,0,1,1),zero3(0,0,1,1),ppf)
      end subroutine ref_steppi

! s_phy2cnt (phy2cnt.f90:193-227): contravariant w.
      subroutine ref_phy2cnt(ni,nj,nk,jcb8w,u,v,w,wc)                    &
     &   bind(C,name='ref_phy2cnt')
      integer(c_int), value :: ni,nj,nk
      real jcb8w(*),u(*),v(*),w(*),wc(*)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_phy2cnt: setup'
      call s_phy2cnt(idsthopt,idtrnopt,idmpopt,idmfcopt,idadvopt,        &
     &  ni,nj,nk,z3,z3,jcb8w,zmf,u,v,w,wc,wk(0,0,1,17),wk(0,0,1,18),     &
     &  wk(0,0,1,19))
      end subroutine ref_phy2cnt

! s_steps (steps.f90:114-4501) on the past/future levels of lev (as in
! ref_intgdrv) with the forcings ptfrc, qvfrc and, as fdmdrv passes them,
! the future q/nci slots holding the species' forcings on entry.
      subroutine ref_steps(ni,nj,nk,moist,dtb,qvbr,rst,lev,ptfrc,qvfrc,  &
     &   dtdrst) bind(C,name='ref_steps')
      integer(c_int), value :: ni,nj,nk,moist
      real, value :: dtb
      real qvbr(*),rst(*),ptfrc(*),qvfrc(*),dtdrst(*)
      real lev(0:ni+1,0:nj+1,1:nk,1:16,1:3)
      character(len=5) fmois
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_steps: setup'
      if(moist.ne.0) then
        fmois='moist'
      else
        fmois='dry  '
      end if
      call s_steps(idgpvvar,idexbvar,idexbopt,idgwmopt,idadvopt,        &
     &  idcphopt,idhaiopt,idqcgopt,idaslopt,idtrkopt,idtubopt,fmois,    &
     &  dtb,0.e0,0.e0,ni,nj,nk,nqw,nnw,nqi,nni,nqa,qvbr,rst,            &
     &  lev(0,0,1,5,2),lev(0,0,1,5,1),lev(0,0,1,6,2),lev(0,0,1,6,1),    &
     &  lev(0,0,1,7,2),lev(0,0,1,7,1),lev(0,0,1,12,2),lev(0,0,1,12,1),  &
     &  lev(0,0,1,9,2),lev(0,0,1,9,1),lev(0,0,1,14,2),lev(0,0,1,14,1),  &
     &  zero3(0,0,1,11),zero3(0,0,1,14),zero3(0,0,1,17),zero3(0,0,1,20),&
     &  zero3(0,0,1,23),zero3(0,0,1,24),zero3(0,0,1,25),zero3(0,0,1,26),&
     &  zero3(0,0,1,27),zero3(0,0,1,28),                                &
     &  cpx(1,1,1,5),cpy(1,1,1,5),cpx(1,1,1,6),cpy(1,1,1,6),            &
     &  cpx(1,1,1,7),cpy(1,1,1,7),cpx(1,1,1,9),cpy(1,1,1,9),            &
     &  cpx(1,1,1,11),cpy(1,1,1,11),cpx(1,1,1,14),cpy(1,1,1,14),        &
     &  cpx(1,1,1,17),cpy(1,1,1,17),cpx(1,1,1,19),cpy(1,1,1,19),        &
     &  cpx(1,1,1,22),cpy(1,1,1,22),cpx(1,1,1,23),cpy(1,1,1,23),        &
     &  cpx(1,1,1,24),cpy(1,1,1,24),                                    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),                                  &
     &  ptfrc,qvfrc,lev(0,0,1,7,3),lev(0,0,1,12,3),lev(0,0,1,9,3),      &
     &  lev(0,0,1,14,3),zero3(0,0,1,29),zero3(0,0,1,32),                &
     &  zero3(0,0,1,35),zero3(0,0,1,36),zero3(0,0,1,37),                &
     &  lev(0,0,1,5,3),lev(0,0,1,6,3),dtdrst)
      end subroutine ref_steps

! s_steps with tubopt=2: as ref_steps, with the TKE levels tke, tkep and
! the TKE forcing / future level tkef in place of the zero arrays.
      subroutine ref_steps_tke(ni,nj,nk,dtb,qvbr,rst,lev,ptfrc,qvfrc,     &
     &   dtdrst,tke,tkep,tkef) bind(C,name='ref_steps_tke')
      integer(c_int), value :: ni,nj,nk
      real, value :: dtb
      real qvbr(*),rst(*),ptfrc(*),qvfrc(*),dtdrst(*),tke(*),tkep(*)
      real tkef(*)
      real lev(0:ni+1,0:nj+1,1:nk,1:16,1:3)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_steps_tke: setup'
      iname(idtubopt)=2
      call s_steps(idgpvvar,idexbvar,idexbopt,idgwmopt,idadvopt,        &
     &  idcphopt,idhaiopt,idqcgopt,idaslopt,idtrkopt,idtubopt,'moist',  &
     &  dtb,0.e0,0.e0,ni,nj,nk,nqw,nnw,nqi,nni,nqa,qvbr,rst,            &
     &  lev(0,0,1,5,2),lev(0,0,1,5,1),lev(0,0,1,6,2),lev(0,0,1,6,1),    &
     &  lev(0,0,1,7,2),lev(0,0,1,7,1),lev(0,0,1,12,2),lev(0,0,1,12,1),  &
     &  lev(0,0,1,9,2),lev(0,0,1,9,1),lev(0,0,1,14,2),lev(0,0,1,14,1),  &
     &  zero3(0,0,1,11),zero3(0,0,1,14),zero3(0,0,1,17),zero3(0,0,1,20),&
     &  zero3(0,0,1,23),zero3(0,0,1,24),zero3(0,0,1,25),zero3(0,0,1,26),&
     &  tke,tkep,                                                       &
     &  cpx(1,1,1,5),cpy(1,1,1,5),cpx(1,1,1,6),cpy(1,1,1,6),            &
     &  cpx(1,1,1,7),cpy(1,1,1,7),cpx(1,1,1,9),cpy(1,1,1,9),            &
     &  cpx(1,1,1,11),cpy(1,1,1,11),cpx(1,1,1,14),cpy(1,1,1,14),        &
     &  cpx(1,1,1,17),cpy(1,1,1,17),cpx(1,1,1,19),cpy(1,1,1,19),        &
     &  cpx(1,1,1,22),cpy(1,1,1,22),cpx(1,1,1,23),cpy(1,1,1,23),        &
     &  cpx(1,1,1,24),cpy(1,1,1,24),                                    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),zero3(0,0,1,1),    &
     &  zero3(0,0,1,1),zero3(0,0,1,1),                                  &
     &  ptfrc,qvfrc,lev(0,0,1,7,3),lev(0,0,1,12,3),lev(0,0,1,9,3),      &
     &  lev(0,0,1,14,3),zero3(0,0,1,29),zero3(0,0,1,32),                &
     &  zero3(0,0,1,35),zero3(0,0,1,36),tkef,                           &
     &  lev(0,0,1,5,3),lev(0,0,1,6,3),dtdrst)
      iname(idtubopt)=0
      end subroutine ref_steps_tke

! s_cloudphy (cloudphy.f90:84-474) with cphopt=2: reads the past level of
! lev, updates its future level and the precipitation in place.
      subroutine ref_cloudphy(ni,nj,nk,dtb,base,wf,lev,qallp,prwtr,price)&
     &   bind(C,name='ref_cloudphy')
      integer(c_int), value :: ni,nj,nk
      real, value :: dtb
      real base(0:ni+1,0:nj+1,1:nk,1:15),wf(*),qallp(*)
      real lev(0:ni+1,0:nj+1,1:nk,1:16,1:3)
      real prwtr(0:ni+1,0:nj+1,1:2,1:2),price(0:ni+1,0:nj+1,1:2,1:3)
      integer nclstp(0:3)
      real dtcl(1:3)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_cloudphy: setup'
      nclstp(:)=0
      dtcl(:)=dtb
      call s_cloudphy(idadvopt,idcphopt,idhaiopt,nclstp,dtb,dtcl,        &
     &  ni,nj,nk,nqw,nnw,nqi,nni,nk,base(0,0,1,1),base(0,0,1,4),        &
     &  base(0,0,1,7),base(0,0,1,8),base(0,0,1,10),base(0,0,1,11),wf,   &
     &  lev(0,0,1,4,1),lev(0,0,1,5,1),lev(0,0,1,6,1),lev(0,0,1,7,1),    &
     &  lev(0,0,1,9,1),qallp,lev(0,0,1,5,3),lev(0,0,1,6,3),             &
     &  lev(0,0,1,7,3),lev(0,0,1,12,3),lev(0,0,1,12,1),lev(0,0,1,9,3),  &
     &  lev(0,0,1,14,3),lev(0,0,1,14,1),zero3(0,0,1,29),                &
     &  zero3(0,0,1,32),prwtr,price,wk(0,0,1,1),wk(0,0,1,2),            &
     &  wk(0,0,1,3),wk(0,0,1,4),wk(0,0,1,5),wk(0,0,1,6),wk(0,0,1,7),    &
     &  wk(0,0,1,8),wk(0,0,1,9),wk(0,0,1,10),wk(0,0,1,11),wk(0,0,1,12), &
     &  wk(0,0,1,13),wk(0,0,1,14))
      end subroutine ref_cloudphy

! s_fallblk (fallblk.f90:87-551): sedimentation of the future level with
! given fall speeds uq(:,:,:,1:5) = ucq,urq,uiq,usq,ugq and
! un(:,:,:,1:5) = ucn,urn,uin,usn,ugn; q(:,:,:,1:5) = qc,qr,qi,qs,qg,
! nc(:,:,:,1:5) = ncc,ncr,nci,ncs,ncg.
      subroutine ref_fallblk(ni,nj,nk,dtb,jcb,rbr,rst,uq,un,q,nc,        &
     &   prwtr,price) bind(C,name='ref_fallblk')
      integer(c_int), value :: ni,nj,nk
      real, value :: dtb
      real jcb(*),rbr(*),rst(*)
      real uq(0:ni+1,0:nj+1,1:nk,1:5),un(0:ni+1,0:nj+1,1:nk,1:5)
      real q(0:ni+1,0:nj+1,1:nk,1:5),nc(0:ni+1,0:nj+1,1:nk,1:5)
      real prwtr(0:ni+1,0:nj+1,1:2,1:2),price(0:ni+1,0:nj+1,1:2,1:3)
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_fallblk: setup'
      call s_fallblk(idcphopt,idhaiopt,idqcgopt,iddz,dtb,ni,nj,nk,jcb,   &
     &  rbr,rst,uq(0,0,1,1),uq(0,0,1,2),uq(0,0,1,3),uq(0,0,1,4),         &
     &  uq(0,0,1,5),uq(0,0,1,5),un(0,0,1,1),un(0,0,1,2),un(0,0,1,3),    &
     &  un(0,0,1,4),un(0,0,1,5),un(0,0,1,5),q(0,0,1,1),q(0,0,1,2),      &
     &  q(0,0,1,3),q(0,0,1,4),q(0,0,1,5),zero3(0,0,1,39),nc(0,0,1,1),   &
     &  nc(0,0,1,2),nc(0,0,1,3),nc(0,0,1,4),nc(0,0,1,5),                &
     &  zero3(0,0,1,40),zero3(0,0,1,29),zero3(0,0,1,30),                &
     &  zero3(0,0,1,32),zero3(0,0,1,33),zero3(0,0,1,34),                &
     &  zero3(0,0,1,40),prwtr(0,0,1,1),prwtr(0,0,1,2),price(0,0,1,1),   &
     &  price(0,0,1,2),price(0,0,1,3),zero2(0,0,12),wk(0,0,1,1))
      end subroutine ref_fallblk

! s_shift2nd (shift2nd.f90), the 2nd halo row of one level of lev.
      subroutine ref_shift2nd(ni,nj,nk,moist,lv,lev)                     &
     &   bind(C,name='ref_shift2nd')
      integer(c_int), value :: ni,nj,nk,moist,lv
      real lev(0:ni+1,0:nj+1,1:nk,1:16,1:3)
      character(len=5) fmois
      if(ni.ne.mi.or.nj.ne.mj.or.nk.ne.mk) stop 'ref_shift2nd: setup'
      if(moist.ne.0) then
        fmois='moist'
      else
        fmois='dry  '
      end if
      call s_shift2nd(idwbc,idebc,idsbc,idnbc,idadvopt,idsmtopt,         &
     &  idcphopt,idhaiopt,idqcgopt,idaslopt,idtrkopt,idtubopt,fmois,    &
     &  'oooooooooo',ni,nj,nk,nqw,nnw,nqi,nni,nqa,                      &
     &  lev(0,0,1,1,lv),lev(0,0,1,2,lv),lev(0,0,1,3,lv),                &
     &  lev(0,0,1,4,lv),lev(0,0,1,5,lv),lev(0,0,1,6,lv),                &
     &  lev(0,0,1,7,lv),lev(0,0,1,12,lv),lev(0,0,1,9,lv),               &
     &  lev(0,0,1,14,lv),zero3(0,0,1,29),zero3(0,0,1,32),               &
     &  zero3(0,0,1,35),zero3(0,0,1,36),zero3(0,0,1,37))
      end subroutine ref_shift2nd

! The reference's base-state setup over flat terrain, in the order the
! solver runs it: s_stretch with ztop1=z(nk-2), ztop2=z(nk-1)
! (phycood.f90:178-179; z(k)=zsfc+(k-2)dz, getz.f90:113; sthopt=0 takes z
! itself), zph = zsth, jcb as s_jacobian forms it over zeta
! (jacobian.f90:243-251) and its averages by s_var8uvw
! (jacobian.f90:339), then s_intrpsnd, s_setbase, s_baserho
! (inivar.f90:667-684), s_sndwave (inivar.f90:850) and s_vspdmp.
! base gets jcb,jcb8u,jcb8v,jcb8w,ubr,vbr,pbr,ptbr,qvbr,rbr,rst,rst8u,
! rst8v,rst8w,rcsq (the harness's base block order).
      subroutine ref_setup_base(ni,nj,nk,sthopt,dz,zsfc,dzmin,layer1,     &
     &   layer2,nlev,z1d,u1d,v1d,p1d,pt1d,qv1d,botbar,vspbar,zsth,base,  &
     &   rbct,ksp0) bind(C,name='ref_setup_base')
      integer(c_int), value :: ni,nj,nk,sthopt,nlev
      real, value :: dz,zsfc,dzmin,layer1,layer2,botbar,vspbar
      real, intent(in) :: z1d(0:nlev),u1d(0:nlev),v1d(0:nlev)
      real, intent(in) :: p1d(0:nlev),pt1d(0:nlev),qv1d(0:nlev)
      real, intent(out) :: zsth(1:nk)
      real, intent(inout) :: base(0:ni+1,0:nj+1,1:nk,1:15)
      real, intent(inout) :: rbct(1:ni,1:nj,1:nk,1:2)
      integer(c_int), intent(out) :: ksp0(2)
      real z(1:nk),dzsth(1:nk),z1dmax(1:nk)
      real, allocatable :: zph(:,:,:),t1(:,:,:),t2(:,:,:),t3(:,:,:)
      integer i,j,k,kk(2)
      allocate(zph(0:ni+1,0:nj+1,1:nk))
      allocate(t1(0:ni+1,0:nj+1,1:nk),t2(0:ni+1,0:nj+1,1:nk))
      allocate(t3(0:ni+1,0:nj+1,1:nk))
      iname(idsthopt)=sthopt
      rname(idzsfc)=zsfc
      rname(iddzmin)=dzmin
      rname(idlayer1)=layer1
      rname(idlayer2)=layer2
      cname(idsndtyp)='zpm'
      iname(idvspopt)=2
      rname(idvspgpv)=0.e0
      rname(idbotgpv)=0.e0
      rname(idvspbar)=vspbar
      rname(idbotbar)=botbar
      do k=1,nk
        z(k)=zsfc+real(k-2)*dz
      end do
      if(sthopt.eq.0) then
        zsth(1:nk)=z(1:nk)
      else
        call s_stretch(idsthopt,idzsfc,iddzmin,idlayer1,idlayer2,       &
     &    z(nk-2),z(nk-1),nk,zsth,dzsth)
      end if
      do k=1,nk
        zph(:,:,k)=zsth(k)
      end do
      base(:,:,:,1)=0.e0
      do k=1,nk-1
        do j=0,nj+1
        do i=0,ni+1
          base(i,j,k,1)=(zph(i,j,k+1)-zph(i,j,k))/(z(k+1)-z(k))
        end do
        end do
      end do
      call s_var8uvw(idwbc,idebc,idexbopt,ni,nj,nk,base(0,0,1,1),         &
     &  base(0,0,1,2),base(0,0,1,3),base(0,0,1,4))
      call s_intrpsnd(idsndtyp,ni,nj,nk,zph,base(0,0,1,5),base(0,0,1,6), &
     &  base(0,0,1,7),base(0,0,1,8),base(0,0,1,9),t1,t2,t3,nlev,z1d,u1d, &
     &  v1d,p1d,pt1d,qv1d)
      call s_setbase(ni,nj,nk,zph,base(0,0,1,5),base(0,0,1,6),           &
     &  base(0,0,1,7),base(0,0,1,8),base(0,0,1,9),base(0,0,1,10),t1,t2,  &
     &  t3)
      call s_baserho(idadvopt,idsmtopt,ni,nj,nk,base(0,0,1,1),           &
     &  base(0,0,1,10),base(0,0,1,11),base(0,0,1,12),base(0,0,1,13),     &
     &  base(0,0,1,14))
      call s_sndwave(ni,nj,nk,base(0,0,1,7),base(0,0,1,15))
      call s_vspdmp(idvspopt,idvspgpv,idvspbar,idbotgpv,idbotbar,kk,ni,  &
     &  nj,nk,zph,rbct,z1dmax)
      ksp0(1)=kk(1)
      ksp0(2)=kk(2)
      deallocate(zph,t1,t2,t3)
      end subroutine ref_setup_base

      end module cress_ref_mpi
