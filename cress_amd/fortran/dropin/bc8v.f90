!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_bc8v
! (SrcOrig/bc8v.f90:72): the south and north fills of a
! v-staggered array run on the GPU through cress_bc8v.  Same
! module and dummy list; the options come from the reference's name table
! (m_getiname); one rank, every side an outer boundary.
!-----------------------------------------------------------------------
      module m_bc8v

      use iso_c_binding
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: bc8v, s_bc8v

      interface bc8v
        module procedure s_bc8v
      end interface

      contains

      subroutine s_bc8v(fpsbc,fpnbc,ni,nj,kmax,var8v)

      integer, intent(in) :: fpsbc
      integer, intent(in) :: fpnbc
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: kmax
      real, intent(inout) :: var8v(0:ni+1,0:nj+1,1:kmax)

      type(cress_config) cfg
      integer blo, bhi
      integer(c_int) rc

      call getiname(fpsbc,blo)
      call getiname(fpnbc,bhi)

      call base_cfg(cfg,ni,nj,max(kmax,4),1,1.e0,1.e0,1.e0)
      cfg%sbc=blo
      cfg%nbc=bhi

      rc=cress_bc8v(cfg,int(kmax,c_int),var8v)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_bc8v: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_bc8v

      end module m_bc8v
