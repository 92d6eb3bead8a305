!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_bc8u
! (SrcOrig/bc8u.f90:72): the west and east fills of a
! u-staggered array run on the GPU through cress_bc8u.  Same
! module and dummy list; the options come from the reference's name table
! (m_getiname); one rank, every side an outer boundary.
!-----------------------------------------------------------------------
      module m_bc8u

      use iso_c_binding
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: bc8u, s_bc8u

      interface bc8u
        module procedure s_bc8u
      end interface

      contains

      subroutine s_bc8u(fpwbc,fpebc,ni,nj,kmax,var8u)

      integer, intent(in) :: fpwbc
      integer, intent(in) :: fpebc
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: kmax
      real, intent(inout) :: var8u(0:ni+1,0:nj+1,1:kmax)

      type(cress_config) cfg
      integer blo, bhi
      integer(c_int) rc

      call getiname(fpwbc,blo)
      call getiname(fpebc,bhi)

      call base_cfg(cfg,ni,nj,max(kmax,4),1,1.e0,1.e0,1.e0)
      cfg%wbc=blo
      cfg%ebc=bhi

      rc=cress_bc8u(cfg,int(kmax,c_int),var8u)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_bc8u: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_bc8u

      end module m_bc8u
