!-----------------------------------------------------------------------
! Drop-in replacement for the reference's module m_turbdrv
! (SrcOrig/turbdrv.f90:80-85): the velocity forcings of the 1.5-order
! TKE closure -- N^2, the deformation tensor and its square, the eddy
! viscosities, the stresses and their divergence -- run on the GPU
! through cress_turbdrv.  Same module and dummy list.  Taken: tubopt=2,
! isoopt=1, trnopt=0, mfcopt=0, sfcopt=0, periodic or open (7) sides;
! zph, j31, j32, the map factors and the work arrays t11, t22, t12, t31,
! t32, tmp1 are not read (nor written).  The options come from the
! reference's name tables (m_getiname, m_getrname, ids of m_comindx).
!-----------------------------------------------------------------------
      module m_turbdrv

      use iso_c_binding
      use m_comindx
      use m_getiname
      use m_getrname
      use m_cress_hip

      implicit none

      private

      public :: turbdrv, s_turbdrv

      interface turbdrv
        module procedure s_turbdrv
      end interface

      contains

      subroutine s_turbdrv(fmois,dtb,ni,nj,nk,                          &
     &                     zph,j31,j32,jcb,jcb8u,jcb8v,jcb8w,           &
     &                     mf,mf8u,mf8v,rmf,rmf8u,rmf8v,pbr,ptbr,rbr,   &
     &                     up,vp,wp,ppp,ptpp,qvp,tkep,qallp,            &
     &                     ufrc,vfrc,wfrc,rkh,rkv,priv,ssq,nsq8w,       &
     &                     t13,t23,t33,t11,t22,t12,t31,t32,tmp1)

      character(len=5), intent(in) :: fmois
      integer, intent(in) :: ni
      integer, intent(in) :: nj
      integer, intent(in) :: nk
      real, intent(in) :: dtb
      real, intent(in) :: zph(0:ni+1,0:nj+1,1:nk)
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
      real, intent(in) :: pbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: rbr(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: up(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: vp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: wp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ppp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: ptpp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qvp(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: tkep(0:ni+1,0:nj+1,1:nk)
      real, intent(in) :: qallp(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: ufrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: vfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: wfrc(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: rkh(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: rkv(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: priv(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: ssq(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: nsq8w(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: t13(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: t23(0:ni+1,0:nj+1,1:nk)
      real, intent(out) :: t33(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: t11(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: t22(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: t12(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: t31(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: t32(0:ni+1,0:nj+1,1:nk)
      real, intent(inout) :: tmp1(0:ni+1,0:nj+1,1:nk)

      type(cress_config) cfg
      integer tubopt, isoopt, trnopt, mfcopt, sfcopt, cphopt
      integer wbc, ebc, sbc, nbc, bbc, tbc, moist
      real dx, dy, dz, thresq
      integer(c_int) rc

      call getiname(idtubopt,tubopt)
      call getiname(idisoopt,isoopt)
      call getiname(idtrnopt,trnopt)
      call getiname(idmfcopt,mfcopt)
      call getiname(idsfcopt,sfcopt)
      call getiname(idcphopt,cphopt)
      call getiname(idwbc,wbc)
      call getiname(idebc,ebc)
      call getiname(idsbc,sbc)
      call getiname(idnbc,nbc)
      call getiname(idbbc,bbc)
      call getiname(idtbc,tbc)
      call getrname(iddx,dx)
      call getrname(iddy,dy)
      call getrname(iddz,dz)
      call getrname(idthresq,thresq)

      if(tubopt.ne.2.or.isoopt.ne.1.or.trnopt.ne.0.or.mfcopt.ne.0.or.   &
     &   sfcopt.ne.0.or.(wbc.ne.-1.and.wbc.ne.7).or.                    &
     &   (ebc.ne.-1.and.ebc.ne.7).or.(sbc.ne.-1.and.sbc.ne.7).or.       &
     &   (nbc.ne.-1.and.nbc.ne.7)) then
        write(6,'(a)') 's_turbdrv: outside the options the GPU path takes'
        stop
      end if

      call base_cfg(cfg,ni,nj,nk,1,dx,dy,dz)
      cfg%cphopt=abs(cphopt)
      cfg%wbc=wbc
      cfg%ebc=ebc
      cfg%sbc=sbc
      cfg%nbc=nbc
      cfg%bbc=bbc
      cfg%tbc=tbc
      cfg%thresq=dble(thresq)
      moist=0
      if(fmois(1:5).eq.'moist') moist=1

      rc=cress_turbdrv(cfg,dtb,moist,jcb,jcb8u,jcb8v,jcb8w,pbr,ptbr,    &
     &                 rbr,up,vp,wp,ppp,ptpp,qvp,tkep,qallp,ufrc,vfrc,  &
     &                 wfrc,rkh,rkv,priv,ssq,nsq8w,t13,t23,t33)

      if(rc.ne.cress_ok) then
        write(6,'(2a)') 's_turbdrv: ',trim(cress_errmsg(rc))
        stop
      end if

      end subroutine s_turbdrv

      end module m_turbdrv
