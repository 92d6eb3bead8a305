!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_advuvw
! (SrcOrig/advuvw.f90:75-78): the advection of u, v, w added to ufrc, vfrc,
! wfrc runs on the GPU through cress_advuvw.  Same module and dummy list;
! hadv, vadv and tmp1..tmp3 are the reference's work arrays and are not
! written.  Options come from the reference's name tables (m_getiname,
! m_getrname, ids of m_comindx); the GPU path takes dx, dy, dz and forms
! 1/dx itself in binary32, as the reference's setup does.  One rank; the
! 4th-order terms stop next to sides flagged by iwest..jnorth, which the
! GPU path takes as open sides (option 7).
!-----------------------------------------------------------------------
      module m_advuvw

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: advuvw, s_advuvw

      interface advuvw
        module procedure s_advuvw
      end interface

      contains

      subroutine s_advuvw(fpadvopt,fpiwest,fpieast,fpjsouth,fpjnorth,   &
     &                    fpdxiv,fpdyiv,fpdziv,ni,nj,nk,rstxu,rstxv,    &
     &                    rstxwc,u,v,w,ufrc,vfrc,wfrc,hadv,vadv,        &
     &                    tmp1,tmp2,tmp3)

      integer, intent(in) :: fpadvopt, fpiwest, fpieast, fpjsouth
      integer, intent(in) :: fpjnorth, fpdxiv, fpdyiv, fpdziv
      integer, intent(in) :: ni, nj, nk
      real, intent(in) :: rstxu(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxwc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: w(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: ufrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: vfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: wfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: hadv(0:ni+1,0:nj+1,1:nk)
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

      call base_cfg(cfg,ni,nj,nk,advopt,dx,dy,dz)
      cfg%wbc=merge(7,-1,iwest.ne.0)
      cfg%ebc=merge(7,-1,ieast.ne.0)
      cfg%sbc=merge(7,-1,jsouth.ne.0)
      cfg%nbc=merge(7,-1,jnorth.ne.0)

      rc=cress_advuvw(cfg,rstxu,rstxv,rstxwc,u,v,w,ufrc,vfrc,wfrc)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_advuvw: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_advuvw

      end module m_advuvw
