!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_cloudphy
! (SrcOrig/cloudphy.f90:84-91): the bulk cold-rain microphysics
! (cphopt=2) on a given future level -- saturation adjustment, the
! source terms, sedimentation and the surface precipitation rate and
! accumulation -- runs on the GPU through cress_cloudphy over 2 dtb, as
! the leapfrog schemes (advopt<=3) step it.  Same module and dummy list.
! Taken: advopt<=3, cphopt=2, haiopt=0, periodic sides; jcb8w, wf,
! qallp, the concentrations of water, the charging arrays, nclstp, dtcl
! and the work arrays are not read.  The options come from the
! reference's name tables (m_getiname).
!-----------------------------------------------------------------------
      module m_cloudphy

      use iso_c_binding
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: cloudphy, s_cloudphy

      interface cloudphy
        module procedure s_cloudphy
      end interface

      contains

      subroutine s_cloudphy(fpadvopt,fpcphopt,fphaiopt,nclstp,          &
     &                      dtb,dtcl,ni,nj,nk,nqw,nnw,nqi,nni,km,       &
     &                      jcb,jcb8w,pbr,ptbr,rbr,rst,wf,ppp,          &
     &                      ptpp,qvp,qwtrp,qicep,qallp,ptpf,qvf,        &
     &                      qwtrf,nwtrf,nwtrp,qicef,nicef,nicep,        &
     &                      qcwtrf,qcicef,prwtr,price,rbv,pi,p,         &
     &                      tmp1,tmp2,tmp3,tmp4,tmp5,tmp6,tmp7,         &
     &                      tmp8,tmp9,tmp10,tmp11)

      integer, intent(in) :: fpadvopt
      integer, intent(in) :: fpcphopt
      integer, intent(in) :: fphaiopt
      integer, intent(in) :: nclstp(0:3)
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      integer, intent(in) :: nqw
      integer, intent(in) :: nnw
      integer, intent(in) :: nqi
      integer, intent(in) :: nni
      integer, intent(in) :: km
      real, intent(in) :: dtb
      real, intent(in) :: dtcl(1:3)
      real, intent(in) :: jcb(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8w(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: pbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: wf(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ppp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptpp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qvp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qwtrp(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(in) :: qicep(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(in) :: qallp(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: ptpf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: qvf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: qwtrf(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(inout) :: nwtrf(0:ni+1,0:nj+1,1:nk,1:nnw)
      real, intent(inout) :: nwtrp(0:ni+1,0:nj+1,1:nk,1:nnw)
      real, intent(inout) :: qicef(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(inout) :: nicef(0:ni+1,0:nj+1,1:nk,1:nni)
      real, intent(inout) :: nicep(0:ni+1,0:nj+1,1:nk,1:nni)
      real, intent(inout) :: qcwtrf(0:ni+1,0:nj+1,1:nk,1:nqw)
      real, intent(inout) :: qcicef(0:ni+1,0:nj+1,1:nk,1:nqi)
      real, intent(inout) :: prwtr(0:ni+1,0:nj+1,1:2,1:nqw)
      real, intent(inout) :: price(0:ni+1,0:nj+1,1:2,1:nqi)
      real, intent(inout) :: rbv(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: pi(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: p(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp3(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp4(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp5(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp6(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp7(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp8(0:ni+1,0:nj+1,1:km)
      real, intent(inout) :: tmp9(0:ni+1,0:nj+1,1:km)
      real, intent(inout) :: tmp10(0:ni+1,0:nj+1,1:km)
      real, intent(inout) :: tmp11(0:ni+1,0:nj+1,1:km)

      type(cress_config) cfg
      integer advopt, cphopt, haiopt
      integer(c_int) rc

      call getiname(fpadvopt,advopt)
      call getiname(fpcphopt,cphopt)
      call getiname(fphaiopt,haiopt)

      if(advopt.gt.3.or.cphopt.ne.2.or.haiopt.ne.0.or.nqw.ne.2.or.      &
     &   nqi.ne.3.or.nni.ne.1) then
        write(6,'(a)') 's_cloudphy: outside the options the GPU path takes'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,advopt,1.e0,1.e0,1.e0)
      cfg%cphopt=cphopt

      rc=cress_cloudphy(cfg,dtb,jcb,pbr,ptbr,rbr,rst,ppp,ptpp,qvp,      &
     &                  qwtrp,qicep,nicep,ptpf,qvf,qwtrf,qicef,nicef,   &
     &                  prwtr,price)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_cloudphy: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_cloudphy

      end module m_cloudphy
