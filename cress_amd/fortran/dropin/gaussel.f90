!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_gaussel
! (SrcOrig/gaussel.f90:37-73): same module, generic and specific names and
! the same dummy list; the impopt=1 Thomas solve runs on the GPU through
! cress_gaussel.  It compiles inside the reference tree, next to the
! reference's m_getiname, with m_cress_hip on the include path.
!-----------------------------------------------------------------------
      module m_gaussel

      use iso_c_binding
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: gaussel, s_gaussel

      interface gaussel
        module procedure s_gaussel
      end interface

      contains

      subroutine s_gaussel(fpimpopt,istr,iend,jstr,jend,kstr,kend,ni,nj,&
     &                     kmax,rr,ss,tt,ff,pv)

      integer, intent(in) :: fpimpopt
      integer, intent(in) :: istr, iend, jstr, jend, kstr, kend
      integer, intent(in) :: ni, nj, kmax
      real, intent(inout) :: rr(0:ni+1,0:nj+1,1:kmax)
      real, intent(inout) :: ss(0:ni+1,0:nj+1,1:kmax)
      real, intent(inout) :: tt(0:ni+1,0:nj+1,1:kmax)
      real, intent(inout) :: ff(0:ni+1,0:nj+1,1:kmax)
      real, intent(inout) :: pv(0:ni+1,0:nj+1,1:kmax)

      integer impopt
      integer(c_int) rc

      call getiname(fpimpopt,impopt)

      if(impopt.ne.1) then
        write(6,'(a)') 's_gaussel: only impopt=1 runs on the GPU'
        stop
      end if

      rc=cress_gaussel(istr,iend,jstr,jend,kstr,kend,ni,nj,kmax,        &
     &                 rr,ss,tt,ff)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_gaussel: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_gaussel

      end module m_gaussel
