!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_bc8w
! (SrcOrig/bc8w.f90:70): the bottom and top fills of a
! w-staggered array run on the GPU through cress_bc8w.  Same
! module and dummy list; the options come from the reference's name table
! (m_getiname); one rank, every side an outer boundary.
!-----------------------------------------------------------------------
      module m_bc8w

      use iso_c_binding
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: bc8w, s_bc8w

      interface bc8w
        module procedure s_bc8w
      end interface

      contains

      subroutine s_bc8w(fpbbc,fptbc,ni,nj,nk,var8w)

      integer, intent(in) :: fpbbc
      integer, intent(in) :: fptbc
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      real, intent(inout) :: var8w(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer blo, bhi
      integer(c_int) rc

      call getiname(fpbbc,blo)
      call getiname(fptbc,bhi)

      call base_cfg(cfg,ni,nj,nk,1,1.e0,1.e0,1.e0)
      cfg%bbc=blo
      cfg%tbc=bhi

      rc=cress_bc8w(cfg,var8w)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_bc8w: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_bc8w

      end module m_bc8w
