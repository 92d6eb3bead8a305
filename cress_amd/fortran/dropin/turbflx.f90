!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_turbflx
! (SrcOrig/turbflx.f90:72-74): the turbulent fluxes of a scalar at u, v
! and w points run on the GPU through cress_turbflx.  Same module and
! dummy list.  Taken: trnopt=0, sfcopt=0 (j31, j32, sfrc and the work
! arrays j31s, j32s, jcbs are not read).  The grid spacings come from the
! reference's name tables (iddx, iddy, iddz of m_comindx); the inverse
! spacings the caller names are formed from them as the reference forms
! them (1/dx in single precision).
!-----------------------------------------------------------------------
      module m_turbflx

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: turbflx, s_turbflx

      interface turbflx
        module procedure s_turbflx
      end interface

      contains

      subroutine s_turbflx(fptrnopt,fpsfcopt,fpdxiv,fpdyiv,fpdziv,      &
     &                     ni,nj,nk,j31,j32,jcb,s,sfrc,rkh8u,rkh8v,     &
     &                     rkv8w,h1,h2,h3,j31s,j32s,jcbs)

      integer, intent(in) :: fptrnopt
      integer, intent(in) :: fpsfcopt
      integer, intent(in) :: fpdxiv
      integer, intent(in) :: fpdyiv
      integer, intent(in) :: fpdziv
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      real, intent(in) :: j31(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: j32(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: s(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: sfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rkh8u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rkh8v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rkv8w(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: h1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: h2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: h3(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: j31s(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: j32s(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: jcbs(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer trnopt, sfcopt
      real dx, dy, dz
      integer(c_int) rc

      call getiname(fptrnopt,trnopt)
      call getiname(fpsfcopt,sfcopt)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)

      if(trnopt.ne.0.or.sfcopt.ne.0.or.fpdxiv.ne.iddxiv.or.             &
     &   fpdyiv.ne.iddyiv.or.fpdziv.ne.iddziv) then
        write(6,'(a)') 's_turbflx: outside the options the GPU path takes'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,1,dx,dy,dz)

      rc=cress_turbflx(cfg,jcb,s,rkh8u,rkh8v,rkv8w,h1,h2,h3)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_turbflx: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_turbflx

      end module m_turbflx
