!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_steps
! (SrcOrig/steps.f90:114-127): the leapfrog of theta', qv and, with the
! cold-rain scheme (cphopt=2), qc, qr, qi, qs, qg and nci, with the halo
! fills and vbcs, runs on the GPU through cress_steps, and with the TKE
! closure (tubopt=2) the TKE's through cress_stepstke.  Same module and
! dummy list.  Taken: gwmopt=0, exbopt=0, advopt<=3, a moist state,
! cphopt 0 or 2 with haiopt=0, no aerosol or tracer, no GPV nudging,
! periodic sides (the phase speeds and nudging arrays are not read).  The
! options come from the reference's name tables (m_getiname, ids of
! m_comindx).
!-----------------------------------------------------------------------
      module m_steps

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: steps, s_steps

      interface steps
        module procedure s_steps
      end interface

      contains

      subroutine s_steps(fpgpvvar,fpexbvar,fpexbopt,fpgwmopt,fpadvopt,  &
     &                   fpcphopt,fphaiopt,fpqcgopt,fpaslopt,fptrkopt,  &
     &                   fptubopt,fmois,dtb,gtinc,atinc,ni,nj,nk,       &
     &                   nqw,nnw,nqi,nni,nqa,qvbr,rst,ptp,ptpp,qv,qvp,  &
     &                   qwtr,qwtrp,nwtr,nwtrp,qice,qicep,nice,nicep,   &
     &                   qcwtr,qcwtrp,qcice,qcicep,qasl,qaslp,qt,qtp,   &
     &                   tke,tkep,ptcpx,ptcpy,qvcpx,qvcpy,qwcpx,qwcpy,  &
     &                   nwcpx,nwcpy,qicpx,qicpy,nicpx,nicpy,           &
     &                   qcwcpx,qcwcpy,qcicpx,qcicpy,qacpx,qacpy,       &
     &                   qtcpx,qtcpy,tkecpx,tkecpy,ptpgpv,ptptd,        &
     &                   qvgpv,qvtd,qwgpv,qwtd,qigpv,qitd,qagpv,qatd,   &
     &                   ptfrc,qvfrc,qwtrf,nwtrf,qicef,nicef,           &
     &                   qcwtrf,qcicef,qaslf,qtf,tkef,                  &
     &                   ptpf,qvf,dtdrst)

      character(len=5), intent(in) :: fmois
      integer, intent(in) :: fpgpvvar
      integer, intent(in) :: fpexbvar
      integer, intent(in) :: fpexbopt
      integer, intent(in) :: fpgwmopt
      integer, intent(in) :: fpadvopt
      integer, intent(in) :: fpcphopt
      integer, intent(in) :: fphaiopt
      integer, intent(in) :: fpqcgopt
      integer, intent(in) :: fpaslopt
      integer, intent(in) :: fptrkopt
      integer, intent(in) :: fptubopt
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      integer, intent(in) :: nqw
      integer, intent(in) :: nnw
      integer, intent(in) :: nqi
      integer, intent(in) :: nni
      integer, intent(in) :: nqa(0:4)
      real, intent(in) :: dtb
      real, intent(in) :: gtinc
      real, intent(in) :: atinc
      real, intent(in) :: qvbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptpp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qvp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qwtr(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: qwtrp(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: nwtr(0:ni+1,0:nj+1,1:nk,1:nnw)
      real, intent(in) :: nwtrp(0:ni+1,0:nj+1,1:nk,1:nnw)
      real, intent(in) :: qice(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: qicep(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: nice(0:ni+1,0:nj+1,1:nk,1:nni)
      real, intent(in) :: nicep(0:ni+1,0:nj+1,1:nk,1:nni)
      real, intent(in) :: qcwtr(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: qcwtrp(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: qcice(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: qcicep(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: qasl(0:ni+1,0:nj+1,1:nk,1:nqa(0))
      real, intent(in) :: qaslp(0:ni+1,0:nj+1,1:nk,1:nqa(0))
      real, intent(in) :: qt(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qtp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: tke(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: tkep(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptcpx(1:nj,1:nk,1:2)
      real, intent(in) :: ptcpy(1:ni,1:nk,1:2)
      real, intent(in) :: qvcpx(1:nj,1:nk,1:2)
      real, intent(in) :: qvcpy(1:ni,1:nk,1:2)
      real, intent(in) :: qwcpx(1:nj,1:nk,1:2,1:nqw)
      real, intent(in) :: qwcpy(1:ni,1:nk,1:2,1:nqw)
      real, intent(in) :: nwcpx(1:nj,1:nk,1:2,1:nnw)
      real, intent(in) :: nwcpy(1:ni,1:nk,1:2,1:nnw)
      real, intent(in) :: qicpx(1:nj,1:nk,1:2,1:nqi)
      real, intent(in) :: qicpy(1:ni,1:nk,1:2,1:nqi)
      real, intent(in) :: nicpx(1:nj,1:nk,1:2,1:nni)
      real, intent(in) :: nicpy(1:ni,1:nk,1:2,1:nni)
      real, intent(in) :: qcwcpx(1:nj,1:nk,1:2,1:nqw)
      real, intent(in) :: qcwcpy(1:ni,1:nk,1:2,1:nqw)
      real, intent(in) :: qcicpx(1:nj,1:nk,1:2,1:nqi)
      real, intent(in) :: qcicpy(1:ni,1:nk,1:2,1:nqi)
      real, intent(in) :: qacpx(1:nj,1:nk,1:2,1:nqa(0))
      real, intent(in) :: qacpy(1:ni,1:nk,1:2,1:nqa(0))
      real, intent(in) :: qtcpx(1:nj,1:nk,1:2)
      real, intent(in) :: qtcpy(1:ni,1:nk,1:2)
      real, intent(in) :: tkecpx(1:nj,1:nk,1:2)
      real, intent(in) :: tkecpy(1:ni,1:nk,1:2)
      real, intent(in) :: ptpgpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptptd(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qvgpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qvtd(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qwgpv(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: qwtd(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: qigpv(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: qitd(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: qagpv(0:ni+1,0:nj+1,1:nk,1:nqa(0))
      real, intent(in) :: qatd(0:ni+1,0:nj+1,1:nk,1:nqa(0))
      real, intent(in) :: ptfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qvfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: qwtrf(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(inout) :: nwtrf(0:ni+1,0:nj+1,1:nk,1:nnw)
      real, intent(inout) :: qicef(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(inout) :: nicef(0:ni+1,0:nj+1,1:nk,1:nni)
      real, intent(inout) :: qcwtrf(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(inout) :: qcicef(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(inout) :: qaslf(0:ni+1,0:nj+1,1:nk,1:nqa(0))
      real, intent(inout) :: qtf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tkef(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: ptpf(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: qvf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: dtdrst(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer gwmopt, exbopt, advopt, cphopt, haiopt, aslopt, trkopt
      integer tubopt, wbc, ebc, sbc, nbc, bbc, tbc
      integer(c_int) rc

      call getiname(fpgwmopt,gwmopt)
      call getiname(fpexbopt,exbopt)
      call getiname(fpadvopt,advopt)
      call getiname(fpcphopt,cphopt)
      call getiname(fphaiopt,haiopt)
      call getiname(fpaslopt,aslopt)
      call getiname(fptrkopt,trkopt)
      call getiname(fptubopt,tubopt)
      call getiname(idwbc,wbc)
      call getiname(idebc,ebc)
      call getiname(idsbc,sbc)
      call getiname(idnbc,nbc)
      call getiname(idbbc,bbc)
      call getiname(idtbc,tbc)

      if(gwmopt.ne.0.or.exbopt.ne.0.or.advopt.gt.3.or.haiopt.ne.0.or.    &
     &   aslopt.ne.0.or.trkopt.ne.0.or.(tubopt.ne.0.and.tubopt.ne.2).or.&
     &   (cphopt.ne.0.and.cphopt.ne.2).or.fmois(1:5).ne.'moist') then
        write(6,'(a)') 's_steps: outside the options the GPU path takes'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,advopt,1.e0,1.e0,1.e0)
      cfg%cphopt=cphopt
      cfg%bbc=bbc
      cfg%tbc=tbc
      if(wbc.ne.-1.or.ebc.ne.-1.or.sbc.ne.-1.or.nbc.ne.-1) then
        cfg%wbc=wbc
      end if

      rc=cress_steps(cfg,dtb,rst,ptpp,ptfrc,qvp,qvfrc,qwtrp,qicep,nicep, &
     &               qwtrf,qicef,nicef,ptpf,qvf,dtdrst)

      if(rc.eq.cress_ok.and.tubopt.ge.2) then
        rc=cress_stepstke(cfg,dtb,rst,tkep,tkef)
      end if

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_steps: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_steps

      end module m_steps
