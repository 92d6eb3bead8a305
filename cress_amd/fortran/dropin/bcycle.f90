!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_bcycle
! (SrcOrig/bcycle.f90:72-74) on one rank: the periodic halo copy of one
! field runs on the GPU through cress_bcycle.  Same module and dummy list;
! compiles inside the reference tree (m_getiname) with m_cress_hip.
!-----------------------------------------------------------------------
      module m_bcycle

      use iso_c_binding
      use m_getiname
      use m_cress_hip

      implicit none

      private

      public :: bcycle, s_bcycle

      interface bcycle
        module procedure s_bcycle
      end interface

      contains

      subroutine s_bcycle(fpwbc,fpebc,fpsbc,fpnbc,                      &
     &                    iwsnd,iwrcv,iesnd,iercv,                      &
     &                    jssnd,jsrcv,jnsnd,jnrcv,ni,nj,kmax,var)

      integer, intent(in) :: fpwbc, fpebc, fpsbc, fpnbc
      integer, intent(in) :: iwsnd, iwrcv, iesnd, iercv
      integer, intent(in) :: jssnd, jsrcv, jnsnd, jnrcv
      integer, intent(in) :: ni, nj, kmax
      real, intent(inout) :: var(0:ni+1,0:nj+1,1:kmax)

      type(cress_config) cfg
      integer wbc, ebc, sbc, nbc
      integer(c_int) rc

      call getiname(fpwbc,wbc)
      call getiname(fpebc,ebc)
      call getiname(fpsbc,sbc)
      call getiname(fpnbc,nbc)

      call dropin_config(ni,nj,kmax,cfg)
      cfg%wbc=wbc
      cfg%ebc=ebc
      cfg%sbc=sbc
      cfg%nbc=nbc

      rc=cress_bcycle(cfg,iwsnd,iwrcv,iesnd,iercv,                      &
     &                jssnd,jsrcv,jnsnd,jnrcv,kmax,var)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_bcycle: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_bcycle

! A configuration record for a single-routine call: dimensions set, every
! option at the plain-path value.
      subroutine dropin_config(ni,nj,nk,cfg)

      integer, intent(in) :: ni, nj, nk
      type(cress_config), intent(out) :: cfg

      cfg%ni=ni
      cfg%nj=nj
      cfg%nk=nk
      cfg%wbc=-1
      cfg%ebc=-1
      cfg%sbc=-1
      cfg%nbc=-1
      cfg%bbc=2
      cfg%tbc=2
      cfg%advopt=3
      cfg%impopt=1
      cfg%gwmopt=0
      cfg%buyopt=1
      cfg%divopt=1
      cfg%cphopt=0
      cfg%haiopt=0
      cfg%vspopt=0
      cfg%coropt=0
      cfg%sthopt=0
      cfg%ksp0(1)=2
      cfg%ksp0(2)=2
      cfg%vspvar(1:10)='x'
      cfg%nsstp=1
      cfg%dx=1.d0
      cfg%dy=1.d0
      cfg%dz=1.d0
      cfg%dtbig=1.d0
      cfg%dtsml=1.d0
      cfg%weicoe=.6d0
      cfg%filcoe=.1d0
      cfg%thresq=1.d-8

      end subroutine dropin_config

      end module m_bcycle
