!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_hevi
! (SrcOrig/hevi.f90:88-95): one acoustic substep of uf, vf, wf, wc, ppf
! runs on the GPU through cress_hevi.  Same module and dummy list.  Taken:
! gwmopt=0 (ptpf is not stepped here, as in the reference for gwmopt=0),
! flat terrain (j31, j32 are not read), no map factors (mf.. rmf8v are not
! read), no GPV nudging (ugpv..ptptd, gtinc are not read), periodic sides
! (the phase speeds are not read); tmp1..tmp6 are not written.  The other
! options come from the reference's name tables (m_getiname, m_getrname,
! ids of m_comindx).
!-----------------------------------------------------------------------
      module m_hevi

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: hevi, s_hevi

      interface hevi
        module procedure s_hevi
      end interface

      contains

      subroutine s_hevi(fpgwmopt,fpbuyopt,isstp,dts,gtinc,              &
     &                  ni,nj,nk,j31,j32,jcb,jcb8u,jcb8v,jcb8w,         &
     &                  mf,mf8u,mf8v,rmf,rmf8u,rmf8v,ubr,vbr,ptbr,rbr,  &
     &                  rst,rst8u,rst8v,rst8w,rcsq,ufrc,vfrc,wfrc,      &
     &                  pfrc,ptfrc,ucpx,ucpy,vcpx,vcpy,wcpx,wcpy,       &
     &                  pcpx,pcpy,ptcpx,ptcpy,ugpv,utd,vgpv,vtd,        &
     &                  wgpv,wtd,ppgpv,pptd,ptpgpv,ptptd,uf,vf,wf,wc,   &
     &                  ppf,ptpf,tmp1,tmp2,tmp3,tmp4,tmp5,tmp6)

      integer, intent(in) :: fpgwmopt
      integer, intent(in) :: fpbuyopt
      integer, intent(in) :: isstp
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      real, intent(in) :: dts
      real, intent(in) :: gtinc
      real, intent(in) :: j31(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: j32(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: jcb8w(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: mf(0:ni+1,0:nj+1)
      real, intent(in) :: mf8u(0:ni+1,0:nj+1)
      real, intent(in) :: mf8v(0:ni+1,0:nj+1)
      real, intent(in) :: rmf(0:ni+1,0:nj+1,1:4)
      real, intent(in) :: rmf8u(0:ni+1,0:nj+1,1:3)
      real, intent(in) :: rmf8v(0:ni+1,0:nj+1,1:3)
      real, intent(in) :: ubr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: vbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst8u(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst8v(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rst8w(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rcsq(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ufrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: vfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: wfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: pfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ucpx(1:nj,1:nk,1:2)
      real, intent(in) :: ucpy(1:ni,1:nk,1:2)
      real, intent(in) :: vcpx(1:nj,1:nk,1:2)
      real, intent(in) :: vcpy(1:ni,1:nk,1:2)
      real, intent(in) :: wcpx(1:nj,1:nk,1:2)
      real, intent(in) :: wcpy(1:ni,1:nk,1:2)
      real, intent(in) :: pcpx(1:nj,1:nk,1:2)
      real, intent(in) :: pcpy(1:ni,1:nk,1:2)
      real, intent(in) :: ptcpx(1:nj,1:nk,1:2)
      real, intent(in) :: ptcpy(1:ni,1:nk,1:2)
      real, intent(in) :: ugpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: utd(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: vgpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: vtd(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: wgpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: wtd(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ppgpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: pptd(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptpgpv(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptptd(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: uf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: vf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: wf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: wc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: ppf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: ptpf(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp2(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp3(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp4(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp5(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp6(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer gwmopt, buyopt, divopt, wbc, ebc, sbc, nbc, bbc, tbc
      integer advopt, sthopt, trnopt, mfcopt
      real dx, dy, dz, weicoe
      integer(c_int) rc

      call getiname(fpgwmopt,gwmopt)
      call getiname(fpbuyopt,buyopt)
      call getiname(iddivopt,divopt)
      call getiname(idwbc,wbc)
      call getiname(idebc,ebc)
      call getiname(idsbc,sbc)
      call getiname(idnbc,nbc)
      call getiname(idbbc,bbc)
      call getiname(idtbc,tbc)
      call getiname(idadvopt,advopt)
      call getiname(idsthopt,sthopt)
      call getiname(idtrnopt,trnopt)
      call getiname(idmfcopt,mfcopt)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)
      call getrname(idweicoe,weicoe)

      if(gwmopt.ne.0.or.trnopt.ne.0.or.mfcopt.ne.0) then
        write(6,'(a)') 's_hevi: the GPU path takes gwmopt=0, trnopt=0, mfcopt=0'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,advopt,dx,dy,dz)
      cfg%bbc=bbc
      cfg%tbc=tbc
      cfg%buyopt=buyopt
      cfg%divopt=divopt
      cfg%sthopt=sthopt
      cfg%weicoe=weicoe
      if(wbc.ne.-1.or.ebc.ne.-1.or.sbc.ne.-1.or.nbc.ne.-1) then
        cfg%wbc=wbc
      end if

      rc=cress_hevi(cfg,dts,jcb,jcb8u,jcb8v,jcb8w,rbr,rst,rst8u,rst8v,  &
     &              rst8w,rcsq,ufrc,vfrc,wfrc,pfrc,uf,vf,wf,wc,ppf)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_hevi: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_hevi

      end module m_hevi
