!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_advp
! (SrcOrig/advp.f90:75-79): the advection of the pressure perturbation
! written into pfrc runs on the GPU through cress_advp.  Same module and
! dummy list; jcbxu..tmp3 are the reference's work arrays and are not
! written.  Map factors (mfcopt) and the diabatic option (diaopt) are not
! taken: both must be 0 (mf8u, mf8v are then not read).  Options come from
! the reference's name tables; the GPU path takes dx, dy, dz and forms 1/dx
! itself in binary32.  One rank; sides flagged by iwest..jnorth are taken
! as open sides (option 7).
!-----------------------------------------------------------------------
      module m_advp

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: advp, s_advp

      interface advp
        module procedure s_advp
      end interface

      contains

      subroutine s_advp(fpadvopt,fpmpopt,fpmfcopt,                      &
     &                  fpdiaopt,fpiwest,fpieast,fpjsouth,fpjnorth,     &
     &                  fpdxiv,fpdyiv,fpdziv,ni,nj,nk,mf8u,mf8v,        &
     &                  jcb8u,jcb8v,jcb8w,u,v,wc,pp,pfrc,               &
     &                  jcbxu,jcbxv,jcbxwc,hadv,vadv,tmp1,tmp2,tmp3)

      integer, intent(in) :: fpadvopt, fpmpopt, fpmfcopt, fpdiaopt
      integer, intent(in) :: fpiwest, fpieast, fpjsouth, fpjnorth
      integer, intent(in) :: fpdxiv, fpdyiv, fpdziv
      integer, intent(in) :: ni, nj, nk
      real, intent(in) :: mf8u(0:ni+1,0:nj+1)
      real, intent(in) :: mf8v(0:ni+1,0:nj+1)
      real, intent(in) :: jcb8u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8w(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: wc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: pp(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: pfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: jcbxu(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: jcbxv(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: jcbxwc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: hadv(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: vadv(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp3(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer advopt, mfcopt, diaopt, iwest, ieast, jsouth, jnorth
      real dx, dy, dz
      integer(c_int) rc

      call getiname(fpadvopt,advopt)
      call getiname(fpmfcopt,mfcopt)
      call getiname(fpdiaopt,diaopt)
      call getiname(fpiwest,iwest)
      call getiname(fpieast,ieast)
      call getiname(fpjsouth,jsouth)
      call getiname(fpjnorth,jnorth)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)

      if(mfcopt.ne.0.or.diaopt.ne.0) then
        write(6,'(a)') 's_advp: the GPU path takes mfcopt=0 and diaopt=0'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,advopt,dx,dy,dz)
      cfg%wbc=merge(7,-1,iwest.ne.0)
      cfg%ebc=merge(7,-1,ieast.ne.0)
      cfg%sbc=merge(7,-1,jsouth.ne.0)
      cfg%nbc=merge(7,-1,jnorth.ne.0)

      rc=cress_advp(cfg,jcb8u,jcb8v,jcb8w,u,v,wc,pp,pfrc)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_advp: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_advp

      end module m_advp
