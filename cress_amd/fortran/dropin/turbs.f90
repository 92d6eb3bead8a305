!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_turbs
! (SrcOrig/turbs.f90:72-75): the divergence of the turbulent fluxes of a
! scalar, added to its forcing, runs on the GPU through cress_turbs.
! Same module and dummy list.  Taken: trnopt=0, mfcopt=0 (j31, j32, the
! map factors and the work arrays are not read).  The grid spacings come
! from the reference's name tables (iddx, iddy, iddz of m_comindx).
!-----------------------------------------------------------------------
      module m_turbs

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: turbs, s_turbs

      interface turbs
        module procedure s_turbs
      end interface

      contains

      subroutine s_turbs(fptrnopt,fpmpopt,fpmfcopt,                     &
     &                   fpdxiv,fpdyiv,fpdziv,ni,nj,nk,j31,j32,         &
     &                   jcb8u,jcb8v,mf,rmf,rmf8u,rmf8v,h1,h2,h3,sfrc,  &
     &                   tmp1,tmp2,tmp3)

      integer, intent(in) :: fptrnopt
      integer, intent(in) :: fpmpopt
      integer, intent(in) :: fpmfcopt
      integer, intent(in) :: fpdxiv
      integer, intent(in) :: fpdyiv
      integer, intent(in) :: fpdziv
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      real, intent(in) :: j31(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: j32(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: mf(0:ni+1,0:nj+1)
      real, intent(in) :: rmf(0:ni+1,0:nj+1,1:4)
      real, intent(in) :: rmf8u(0:ni+1,0:nj+1,1:3)
      real, intent(in) :: rmf8v(0:ni+1,0:nj+1,1:3)
      real, intent(in) :: h1(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: h2(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: h3(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: sfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp3(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer trnopt, mfcopt
      real dx, dy, dz
      integer(c_int) rc

      call getiname(fptrnopt,trnopt)
      call getiname(fpmfcopt,mfcopt)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)

      if(trnopt.ne.0.or.mfcopt.ne.0.or.fpdxiv.ne.iddxiv.or.             &
     &   fpdyiv.ne.iddyiv.or.fpdziv.ne.iddziv) then
        write(6,'(a)') 's_turbs: outside the options the GPU path takes'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,1,dx,dy,dz)

      rc=cress_turbs(cfg,jcb8u,jcb8v,h1,h2,h3,sfrc)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_turbs: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_turbs

      end module m_turbs
