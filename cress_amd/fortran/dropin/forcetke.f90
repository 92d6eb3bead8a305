!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_forcetke
! (SrcOrig/forcetke.f90:88-94): the forcing of the TKE equation --
! advection, shear production, dissipation, diffusion, the vertical
! sponge, buoyancy -- and the scalar diffusivities it leaves in rkh, rkv
! and rkv8s, on the GPU through cress_forcetke.  Same module and dummy
! list.  Taken: tubopt=2, isoopt=1, trnopt=0, mfcopt=0, smtopt=0, no
! lateral sponge on TKE, periodic or open (7) sides; j31, j32, the map
! factors, rbr, rbcxy, h3 and tmp1..tmp3 are not read (nor written), and
! ssq is read only.  The options come from the reference's name tables
! (m_getiname, m_getrname, m_getcname, ids of m_comindx).
!-----------------------------------------------------------------------
      module m_forcetke

      use iso_c_binding
      use m_comindx
      use m_getcname
      use m_getiname
      use m_getrname
      use m_inichar
      use m_cress_hip

      implicit none

      private

      public :: forcetke, s_forcetke

      interface forcetke
        module procedure s_forcetke
      end interface

      contains

      subroutine s_forcetke(fplspvar,fpvspvar,                          &
     &                      fplspopt,fpvspopt,fpsmtopt,ksp0,            &
     &                      ni,nj,nk,j31,j32,jcb,jcb8u,jcb8v,jcb8w,mf,  &
     &                      rmf,rmf8u,rmf8v,rbr,rst,rstxu,rstxv,rstxwc, &
     &                      tke,tkep,rbcxy,rbct,priv,nsq8w,ssq,rkh,rkv, &
     &                      tkefrc,h3,rkv8s,tmp1,tmp2,tmp3)

      integer, intent(in) :: fplspvar
      integer, intent(in) :: fpvspvar
      integer, intent(in) :: fplspopt
      integer, intent(in) :: fpvspopt
      integer, intent(in) :: fpsmtopt
      integer, intent(in) :: ksp0(1:2)
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      real, intent(in) :: j31(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: j32(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8w(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: mf(0:ni+1,0:nj+1)
      real, intent(in) :: rmf(0:ni+1,0:nj+1,1:4)
      real, intent(in) :: rmf8u(0:ni+1,0:nj+1,1:3)
      real, intent(in) :: rmf8v(0:ni+1,0:nj+1,1:3)
      real, intent(in) :: rbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxu(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rstxwc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: tke(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: tkep(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rbcxy(1:ni,1:nj)
      real, intent(in) :: rbct(1:ni,1:nj,1:nk,1:2)
      real, intent(in) :: priv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: nsq8w(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: ssq(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: rkh(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: rkv(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: tkefrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: h3(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: rkv8s(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp3(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      character(len=108) lspvar, vspvar
      integer lspopt, vspopt, smtopt, advopt, tubopt, isoopt, trnopt
      integer mfcopt, wbc, ebc, sbc, nbc, iwest, ieast, jsouth, jnorth
      real dx, dy, dz
      integer(c_int) rc

      call inichar(lspvar)
      call inichar(vspvar)
      call getcname(fplspvar,lspvar)
      call getcname(fpvspvar,vspvar)
      call getiname(fplspopt,lspopt)
      call getiname(fpvspopt,vspopt)
      call getiname(fpsmtopt,smtopt)
      call getiname(idadvopt,advopt)
      call getiname(idtubopt,tubopt)
      call getiname(idisoopt,isoopt)
      call getiname(idtrnopt,trnopt)
      call getiname(idmfcopt,mfcopt)
      call getiname(idwbc,wbc)
      call getiname(idebc,ebc)
      call getiname(idsbc,sbc)
      call getiname(idnbc,nbc)
      call getiname(idiwest,iwest)
      call getiname(idieast,ieast)
      call getiname(idjsouth,jsouth)
      call getiname(idjnorth,jnorth)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)

! the 4th-order reduction flags of one rank follow the sides
! (sethalo.f90:536-580): 0 on periodic sides, 1 on open ones
      if(tubopt.ne.2.or.isoopt.ne.1.or.trnopt.ne.0.or.mfcopt.ne.0.or.   &
     &   mod(smtopt,10).ne.0.or.                                        &
     &   (lspopt.ge.1.and.lspvar(10:10).eq.'o').or.                     &
     &   (wbc.ne.-1.and.wbc.ne.7).or.(ebc.ne.-1.and.ebc.ne.7).or.       &
     &   (sbc.ne.-1.and.sbc.ne.7).or.(nbc.ne.-1.and.nbc.ne.7).or.       &
     &   iwest.ne.merge(0,1,wbc.eq.-1).or.ieast.ne.merge(0,1,ebc.eq.-1) &
     &   .or.jsouth.ne.merge(0,1,sbc.eq.-1).or.                         &
     &   jnorth.ne.merge(0,1,nbc.eq.-1)) then
        write(6,'(a)') 's_forcetke: outside the options the GPU path takes'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,advopt,dx,dy,dz)
      cfg%wbc=wbc
      cfg%ebc=ebc
      cfg%sbc=sbc
      cfg%nbc=nbc
      cfg%vspopt=vspopt
      cfg%vspvar(10)=vspvar(10:10)
      cfg%ksp0(1)=ksp0(1)
      cfg%ksp0(2)=ksp0(2)

      rc=cress_forcetke(cfg,jcb,jcb8u,jcb8v,jcb8w,rst,rstxu,rstxv,      &
     &                  rstxwc,tke,tkep,rbct,priv,nsq8w,ssq,rkh,rkv,    &
     &                  tkefrc,rkv8s)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_forcetke: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_forcetke

      end module m_forcetke
