!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_advs
! (SrcOrig/advs.f90:75-77): scalar advection sfrc of s by the flux weights
! rstxu, rstxv, rstxwc runs on the GPU through cress_advs.  Same module and
! dummy list; vadv and tmp1..tmp3 are the reference's work arrays and are
! not written.  Options come from the reference's name tables (m_getiname,
! m_getrname, ids of m_comindx); the GPU path takes the grid spacings dx,
! dy, dz and forms 1/dx itself in binary32, as the reference's setup does.
! One rank, no lateral 4th-order reduction (iwest=ieast=jsouth=jnorth=0).
!-----------------------------------------------------------------------
      module m_advs

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: advs, s_advs

      interface advs
        module procedure s_advs
      end interface

      contains

      subroutine s_advs(fpadvopt,fpiwest,fpieast,fpjsouth,fpjnorth,     &
     &                  fpdxiv,fpdyiv,fpdziv,ni,nj,nk,rstxu,rstxv,      &
     &                  rstxwc,s,sfrc,vadv,tmp1,tmp2,tmp3)

      integer, intent(in) :: fpadvopt, fpiwest, fpieast, fpjsouth
      integer, intent(in) :: fpjnorth, fpdxiv, fpdyiv, fpdziv
      integer, intent(in) :: ni, nj, nk
      real, intent(in) :: rstxu(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxwc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: s(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: sfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: vadv(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp3(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer advopt, iwest, ieast, jsouth, jnorth
      real dx, dy, dz
      integer(c_int) rc

      call getiname(fpadvopt,advopt)
      call getiname(fpiwest,iwest)
      call getiname(fpieast,ieast)
      call getiname(fpjsouth,jsouth)
      call getiname(fpjnorth,jnorth)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)

      if(iwest.ne.0.or.ieast.ne.0.or.jsouth.ne.0.or.jnorth.ne.0) then
        write(6,'(a)') 's_advs: the GPU path runs one periodic rank'
        stop
      end if

      cfg%ni=ni
      cfg%nj=nj
      cfg%nk=nk
      cfg%wbc=-1
      cfg%ebc=-1
      cfg%sbc=-1
      cfg%nbc=-1
      cfg%bbc=2
      cfg%tbc=2
      cfg%advopt=advopt
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
      cfg%dx=dx
      cfg%dy=dy
      cfg%dz=dz
      cfg%dtbig=1.d0
      cfg%dtsml=1.d0
      cfg%weicoe=.6d0
      cfg%filcoe=.1d0
      cfg%thresq=1.d-8

      rc=cress_advs(cfg,rstxu,rstxv,rstxwc,s,sfrc)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_advs: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_advs

      end module m_advs
