! Test-infrastructure harness: bind(C) entry points that call the reference
! CReSS module procedures unchanged, so that the C restatement in
! oracle/cress_oracle.c can be checked against the reference's own compiled
! code on identical inputs.  Linked only into oracle/_ref/libcressref32.so.
!
! Every array argument is the reference's explicit-shape
! real(0:ni+1,0:nj+1,1:nk) block (column-major, x fastest), passed by address.
! Options are written into the reference's namelist tables (m_comname) with
! the reference's own ids (m_comindx), exactly where getiname/getrname read
! them (getiname.f90:84).

      module cress_ref_harness

      use iso_c_binding
      use m_comindx
      use m_comname
      use m_commpi
      use m_gaussel
      use m_advs
      use m_advuvw
      use m_advp
      use m_pgrad
      use m_diverpih
      use m_diverpiv
      use m_advbspi
      use m_pgradiv
      use m_buoywsi
      use m_bcycle
      use m_buoywb
      use m_coriuv
      use m_rstuvwc
      use m_vspuvw
      use m_forcep
      use m_forcept
      use m_forceqv
      use m_forceq
      use m_timeflt
      use m_totalqwi
      use m_vbcw
      use m_vbcp
      use m_vbcs
      use m_vbcu
      use m_vbcv
      use m_vbcwc
      use m_bbcw
      use m_setref
      use m_getvdens
      use m_getexner
      use m_diagnw
      use m_diagnsg
      use m_adjstnci
      use m_termblk
      use m_srcblk
      use m_upwqp
      use m_upwnp
      use m_satadjst
      use m_phvuvw
      use m_phvs
      use m_rbcu
      use m_rbcv
      use m_rbcw
      use m_rbcss
      use m_rbcpt
      use m_rbcqv
      use m_rbcq
      use m_rbcs
      use m_bc4news
      use m_bc8u
      use m_bc8v
      use m_bc8w
      use m_lbcu
      use m_lbcv
      use m_lbcw
      use m_lbcs
      use m_lbcwc
      use m_smoo4s
      use m_smoo4qv
      use m_smoo4p
      use m_smoo4uvw
      use m_bruntv
      use m_defomten
      use m_defomssq
      use m_eddyvis
      use m_turbdrv
      use m_turbflx
      use m_turbs
      use m_kh8uv
      use m_forcetke

      implicit none

      contains

! Single-rank periodic/closed domain descriptors (the values inimpi/currpe
! give at numpe=1: currpe.f90:105, inimpi_multi.f90:179).
      subroutine ref_configure(wbc,ebc,sbc,nbc,bbc,tbc,advopt,impopt,    &
     &   gwmopt,buyopt,divopt,cphopt,haiopt,vspopt,coropt,               &
     &   dx,dy,dz,weicoe,filcoe,thresq,vspvar10) bind(C,name='ref_configure')
      integer(c_int), value :: wbc,ebc,sbc,nbc,bbc,tbc,advopt,impopt
      integer(c_int), value :: gwmopt,buyopt,divopt,cphopt,haiopt
      integer(c_int), value :: vspopt,coropt
      real, value :: dx,dy,dz,weicoe,filcoe,thresq
      character(kind=c_char), intent(in) :: vspvar10(10)
      integer i
      iname(:)=0
      rname(:)=0.e0
      do i=1,ncn
        cname(i)=repeat('x',108)
      end do
      iname(idwbc)=wbc
      iname(idebc)=ebc
      iname(idsbc)=sbc
      iname(idnbc)=nbc
      iname(idbbc)=bbc
      iname(idtbc)=tbc
      iname(idadvopt)=advopt
      iname(idimpopt)=impopt
      iname(idgwmopt)=gwmopt
      iname(idbuyopt)=buyopt
      iname(iddivopt)=divopt
      iname(idtrnopt)=0
      iname(idmpopt)=4
      iname(idmfcopt)=0
      iname(idsthopt)=1
      iname(idtubopt)=0
      iname(idsmtopt)=0
      iname(idcphopt)=cphopt
      iname(idhaiopt)=haiopt
      iname(idqcgopt)=0
      iname(idvspopt)=vspopt
      iname(idlspopt)=0
      iname(idnggopt)=0
      iname(idexbopt)=0
      iname(idcoropt)=coropt
      iname(idcrvopt)=0
      iname(idsfcopt)=0
      iname(idaslopt)=0
      iname(idtrkopt)=0
      iname(iddiaopt)=0
      iname(idngropt)=0
! 4th-order reduction flags: 0 on periodic sides (sethalo.f90:536-580).
      iname(idiwest)=merge(0,1,wbc.eq.-1)
      iname(idieast)=merge(0,1,ebc.eq.-1)
      iname(idjsouth)=merge(0,1,sbc.eq.-1)
      iname(idjnorth)=merge(0,1,nbc.eq.-1)
      rname(iddx)=dx
      rname(iddy)=dy
      rname(iddz)=dz
      rname(iddxiv)=1.e0/dx
      rname(iddyiv)=1.e0/dy
      rname(iddziv)=1.e0/dz
      rname(idweicoe)=weicoe
      rname(idfilcoe)=filcoe
      rname(idthresq)=thresq
      do i=1,10
        cname(idvspvar)(i:i)=vspvar10(i)
      end do
      nisub=1
      njsub=1
      isub=0
      jsub=0
      nigrp=1
      njgrp=1
      igrp=0
      jgrp=0
      ebw=1
      ebe=1
      ebs=1
      ebn=1
      ebsw=1
      ebse=1
      ebnw=1
      ebne=1
      end subroutine ref_configure

      subroutine ref_gaussel(istr,iend,jstr,jend,kstr,kend,ni,nj,nk,     &
     &   rr,ss,tt,ff,pv) bind(C,name='ref_gaussel')
      integer(c_int), value :: istr,iend,jstr,jend,kstr,kend,ni,nj,nk
      real rr(*),ss(*),tt(*),ff(*),pv(*)
      call s_gaussel(idimpopt,istr,iend,jstr,jend,kstr,kend,ni,nj,nk,    &
     &               rr,ss,tt,ff,pv)
      end subroutine ref_gaussel

      subroutine ref_advs(ni,nj,nk,rstxu,rstxv,rstxwc,s,sfrc,vadv,       &
     &   t1,t2,t3) bind(C,name='ref_advs')
      integer(c_int), value :: ni,nj,nk
      real rstxu(*),rstxv(*),rstxwc(*),s(*),sfrc(*),vadv(*)
      real t1(*),t2(*),t3(*)
      call s_advs(idadvopt,idiwest,idieast,idjsouth,idjnorth,            &
     &            iddxiv,iddyiv,iddziv,ni,nj,nk,rstxu,rstxv,rstxwc,s,     &
     &            sfrc,vadv,t1,t2,t3)
      end subroutine ref_advs

      subroutine ref_advuvw(ni,nj,nk,rstxu,rstxv,rstxwc,u,v,w,ufrc,vfrc, &
     &   wfrc,hadv,vadv,t1,t2,t3) bind(C,name='ref_advuvw')
      integer(c_int), value :: ni,nj,nk
      real rstxu(*),rstxv(*),rstxwc(*),u(*),v(*),w(*)
      real ufrc(*),vfrc(*),wfrc(*),hadv(*),vadv(*),t1(*),t2(*),t3(*)
      call s_advuvw(idadvopt,idiwest,idieast,idjsouth,idjnorth,          &
     &              iddxiv,iddyiv,iddziv,ni,nj,nk,rstxu,rstxv,rstxwc,     &
     &              u,v,w,ufrc,vfrc,wfrc,hadv,vadv,t1,t2,t3)
      end subroutine ref_advuvw

      subroutine ref_rstuvwc(ni,nj,nk,mf8u,mf8v,rst8u,rst8v,rst8w,u,v,wc,&
     &   rstxu,rstxv,rstxwc) bind(C,name='ref_rstuvwc')
      integer(c_int), value :: ni,nj,nk
      real mf8u(*),mf8v(*),rst8u(*),rst8v(*),rst8w(*),u(*),v(*),wc(*)
      real rstxu(*),rstxv(*),rstxwc(*)
      call s_rstuvwc(idmpopt,idmfcopt,idiwest,idieast,idjsouth,idjnorth, &
     &               ni,nj,nk,mf8u,mf8v,rst8u,rst8v,rst8w,u,v,wc,        &
     &               rstxu,rstxv,rstxwc)
      end subroutine ref_rstuvwc

      subroutine ref_buoywb(ni,nj,nk,ptbr,qvbr,rst,ptp,qv,qall,wfrc,qvd, &
     &   wb8s) bind(C,name='ref_buoywb')
      integer(c_int), value :: ni,nj,nk
      real ptbr(*),qvbr(*),rst(*),ptp(*),qv(*),qall(*),wfrc(*),qvd(*)
      real wb8s(*)
      call s_buoywb(idgwmopt,idcphopt,'moist',ni,nj,nk,ptbr,qvbr,rst,    &
     &              ptp,qv,qall,wfrc,qvd,wb8s)
      end subroutine ref_buoywb

      subroutine ref_coriuv(ni,nj,nk,fc,rst,u,v,ufrc,vfrc,t1)            &
     &   bind(C,name='ref_coriuv')
      integer(c_int), value :: ni,nj,nk
      real fc(*),rst(*),u(*),v(*),ufrc(*),vfrc(*),t1(*)
      call s_coriuv(ni,nj,nk,fc,rst,u,v,ufrc,vfrc,t1)
      end subroutine ref_coriuv

      subroutine ref_vspuvw(ksp0,ni,nj,nk,ubr,vbr,rst8u,rst8v,rst8w,     &
     &   up,vp,wp,rbct,dmy,ufrc,vfrc,wfrc,rbct8s) bind(C,name='ref_vspuvw')
      integer(c_int) :: ksp0(2)
      integer(c_int), value :: ni,nj,nk
      real ubr(*),vbr(*),rst8u(*),rst8v(*),rst8w(*),up(*),vp(*),wp(*)
      real rbct(*),dmy(*),ufrc(*),vfrc(*),wfrc(*),rbct8s(*)
      integer k2(2)
      k2(1)=ksp0(1)
      k2(2)=ksp0(2)
      call s_vspuvw(idgpvvar,idvspvar,idvspopt,k2,0.e0,ni,nj,nk,         &
     &              ubr,vbr,rst8u,rst8v,rst8w,up,vp,wp,rbct,             &
     &              dmy,dmy,dmy,dmy,dmy,dmy,ufrc,vfrc,wfrc,rbct8s)
      end subroutine ref_vspuvw

      subroutine ref_forcep(ksp0,ni,nj,nk,jcb,jcb8u,jcb8v,jcb8w,u,v,wc,  &
     &   pp,ppp,rbct,dmy,pfrc,t1,t2,t3,t4,t5,t6,t7,t8)                   &
     &   bind(C,name='ref_forcep')
      integer(c_int) :: ksp0(2)
      integer(c_int), value :: ni,nj,nk
      real jcb(*),jcb8u(*),jcb8v(*),jcb8w(*),u(*),v(*),wc(*),pp(*),ppp(*)
      real rbct(*),dmy(*),pfrc(*),t1(*),t2(*),t3(*),t4(*),t5(*),t6(*)
      real t7(*),t8(*)
      integer k2(2)
      k2(1)=ksp0(1)
      k2(2)=ksp0(2)
      call s_forcep(idnggvar,idlspvar,idvspvar,idlspopt,idvspopt,        &
     &              idsmtopt,k2,0.e0,0.e0,ni,nj,nk,jcb,jcb8u,jcb8v,jcb8w, &
     &              dmy,dmy,u,v,wc,pp,ppp,dmy,rbct,dmy,dmy,pfrc,         &
     &              t1,t2,t3,t4,t5,t6,t7,t8)
      end subroutine ref_forcep

      subroutine ref_forcept(ksp0,ni,nj,nk,jcb,ptbr,rbr,rst,rstxu,rstxv, &
     &   rstxwc,w,ptp,ptpp,rbct,dmy,ptfrc,t1,t2,t3,t4,t5,t6,t7)          &
     &   bind(C,name='ref_forcept')
      integer(c_int) :: ksp0(2)
      integer(c_int), value :: ni,nj,nk
      real jcb(*),ptbr(*),rbr(*),rst(*),rstxu(*),rstxv(*),rstxwc(*)
      real w(*),ptp(*),ptpp(*),rbct(*),dmy(*),ptfrc(*)
      real t1(*),t2(*),t3(*),t4(*),t5(*),t6(*),t7(*)
      integer k2(2)
      k2(1)=ksp0(1)
      k2(2)=ksp0(2)
      call s_forcept(idnggvar,idlspvar,idvspvar,idlspopt,idvspopt,       &
     &               idgwmopt,idsmtopt,idadvopt,idtubopt,k2,0.e0,0.e0,   &
     &               ni,nj,nk,dmy,dmy,jcb,dmy,dmy,dmy,dmy,dmy,dmy,       &
     &               ptbr,rbr,rst,rstxu,rstxv,rstxwc,w,w,ptp,ptpp,       &
     &               dmy,dmy,dmy,dmy,rbct,dmy,dmy,ptfrc,                 &
     &               t1,t2,t3,t4,t5,t6,t7)
      end subroutine ref_forcept

      subroutine ref_forceqv(ksp0,ni,nj,nk,jcb,qvbr,rbr,rst,rstxu,rstxv, &
     &   rstxwc,qv,qvp,rbct,dmy,qvfrc,t1,t2,t3,t4,t5,t6)                 &
     &   bind(C,name='ref_forceqv')
      integer(c_int) :: ksp0(2)
      integer(c_int), value :: ni,nj,nk
      real jcb(*),qvbr(*),rbr(*),rst(*),rstxu(*),rstxv(*),rstxwc(*)
      real qv(*),qvp(*),rbct(*),dmy(*),qvfrc(*)
      real t1(*),t2(*),t3(*),t4(*),t5(*),t6(*)
      integer k2(2)
      k2(1)=ksp0(1)
      k2(2)=ksp0(2)
      call s_forceqv(idnggvar,idlspvar,idvspvar,idlspopt,idvspopt,       &
     &               idsmtopt,idtubopt,k2,0.e0,0.e0,ni,nj,nk,            &
     &               dmy,dmy,jcb,dmy,dmy,dmy,dmy,dmy,dmy,qvbr,rbr,       &
     &               rst,rstxu,rstxv,rstxwc,qv,qvp,dmy,dmy,dmy,          &
     &               dmy,rbct,dmy,dmy,qvfrc,t1,t2,t3,t4,t5,t6)
      end subroutine ref_forceqv

! Hydrometeor forcing for cphopt=2: qwtr(:,:,:,1:2)=qc,qr; qice(1:3)=qi,qs,qg;
! nice(1)=nci.  Arrays are stacked species-last like the reference.
      subroutine ref_forceq(ksp0,ni,nj,nk,jcb,rbr,rst,rstxu,rstxv,rstxwc,&
     &   qwtr,qwtrp,qice,qicep,nice,nicep,rbct,dmy,qwfrc,qifrc,nifrc,    &
     &   t1,t2,t3,t4,t5,t6) bind(C,name='ref_forceq')
      integer(c_int) :: ksp0(2)
      integer(c_int), value :: ni,nj,nk
      real jcb(*),rbr(*),rst(*),rstxu(*),rstxv(*),rstxwc(*)
      real qwtr(*),qwtrp(*),qice(*),qicep(*),nice(*),nicep(*)
      real rbct(*),dmy(*),qwfrc(*),qifrc(*),nifrc(*)
      real t1(*),t2(*),t3(*),t4(*),t5(*),t6(*)
      integer k2(2)
      real ngrdmp(2),rtinc(2)
      k2(1)=ksp0(1)
      k2(2)=ksp0(2)
      ngrdmp(:)=0.e0
      rtinc(:)=0.e0
      call s_forceq(idnggvar,idlspvar,idvspvar,idngrvar,                 &
     &              idlspopt,idvspopt,idngropt,idsmtopt,idcphopt,        &
     &              idhaiopt,idtubopt,k2,0.e0,ngrdmp,0.e0,rtinc,         &
     &              ni,nj,nk,2,1,3,1,dmy,dmy,jcb,dmy,dmy,                &
     &              dmy,dmy,dmy,dmy,rbr,rst,rstxu,rstxv,rstxwc,          &
     &              qwtr,qwtrp,dmy,dmy,qice,qicep,nice,nicep,            &
     &              dmy,dmy,dmy,dmy,rbct,dmy,dmy,dmy,dmy,                &
     &              dmy,dmy,dmy,dmy,qwfrc,dmy,qifrc,nifrc,               &
     &              t1,t2,t3,t4,t5,t6)
      end subroutine ref_forceq

      subroutine ref_pgrad(dts,ni,nj,nk,jcb,rst8u,rst8v,rst8w,u,v,wc,pp, &
     &   upg,vpg,wpg,t1,t2,t3,dmy) bind(C,name='ref_pgrad')
      real, value :: dts
      integer(c_int), value :: ni,nj,nk
      real jcb(*),rst8u(*),rst8v(*),rst8w(*),u(*),v(*),wc(*),pp(*)
      real upg(*),vpg(*),wpg(*),t1(*),t2(*),t3(*),dmy(*)
      call s_pgrad(idtrnopt,idmpopt,idmfcopt,iddivopt,iddx,iddy,iddz,    &
     &             iddxiv,iddyiv,iddziv,dts,ni,nj,nk,dmy,dmy,jcb,        &
     &             dmy,dmy,dmy,dmy,dmy,dmy,rst8u,rst8v,rst8w,u,v,wc,pp,  &
     &             upg,vpg,wpg,t1,t2,t3)
      end subroutine ref_pgrad

      subroutine ref_diverpih(ni,nj,nk,jcb8u,jcb8v,rcsq,u,v,pdiv,t1,t2,  &
     &   t3,dmy) bind(C,name='ref_diverpih')
      integer(c_int), value :: ni,nj,nk
      real jcb8u(*),jcb8v(*),rcsq(*),u(*),v(*),pdiv(*),t1(*),t2(*),t3(*)
      real dmy(*)
      call s_diverpih(idtrnopt,idmpopt,idmfcopt,iddziv,ni,nj,nk,         &
     &                dmy,dmy,jcb8u,jcb8v,dmy,dmy,dmy,dmy,rcsq,u,v,pdiv, &
     &                t1,t2,t3)
      end subroutine ref_diverpih

      subroutine ref_diverpiv(ni,nj,nk,rcsq,w,pdiv)                      &
     &   bind(C,name='ref_diverpiv')
      integer(c_int), value :: ni,nj,nk
      real rcsq(*),w(*),pdiv(*)
      call s_diverpiv(iddziv,ni,nj,nk,rcsq,w,pdiv)
      end subroutine ref_diverpiv

      subroutine ref_advbspi(fore,ni,nj,nk,rst,w,pfrc,phdiv,pvdiv,fp)    &
     &   bind(C,name='ref_advbspi')
      integer(c_int), value :: fore,ni,nj,nk
      real rst(*),w(*),pfrc(*),phdiv(*),pvdiv(*),fp(*)
      if(fore.ne.0) then
        call s_advbspi(idweicoe,'fore',ni,nj,nk,rst,w,pfrc,phdiv,pvdiv,fp)
      else
        call s_advbspi(idweicoe,'back',ni,nj,nk,rst,w,pfrc,phdiv,pvdiv,fp)
      end if
      end subroutine ref_advbspi

      subroutine ref_pgradiv(dts,ni,nj,nk,jcb,wfrc,fp,fw,fpdvj)          &
     &   bind(C,name='ref_pgradiv')
      real, value :: dts
      integer(c_int), value :: ni,nj,nk
      real jcb(*),wfrc(*),fp(*),fw(*),fpdvj(*)
      call s_pgradiv(iddziv,idweicoe,dts,ni,nj,nk,jcb,wfrc,fp,fw,fpdvj)
      end subroutine ref_pgradiv

      subroutine ref_buoywsi(dts,ni,nj,nk,rbr,ptbr,rst,rcsq,pp,ptp,fp,fw,&
     &   wb8s) bind(C,name='ref_buoywsi')
      real, value :: dts
      integer(c_int), value :: ni,nj,nk
      real rbr(*),ptbr(*),rst(*),rcsq(*),pp(*),ptp(*),fp(*),fw(*),wb8s(*)
      call s_buoywsi(idgwmopt,idweicoe,dts,ni,nj,nk,rbr,ptbr,rst,rcsq,   &
     &               pp,ptp,fp,fw,wb8s)
      end subroutine ref_buoywsi

      subroutine ref_bcycle(iwsnd,iwrcv,iesnd,iercv,jssnd,jsrcv,jnsnd,   &
     &   jnrcv,ni,nj,kmax,var) bind(C,name='ref_bcycle')
      integer(c_int), value :: iwsnd,iwrcv,iesnd,iercv
      integer(c_int), value :: jssnd,jsrcv,jnsnd,jnrcv,ni,nj,kmax
      real var(*)
      call s_bcycle(idwbc,idebc,idsbc,idnbc,iwsnd,iwrcv,iesnd,iercv,     &
     &              jssnd,jsrcv,jnsnd,jnrcv,ni,nj,kmax,var)
      end subroutine ref_bcycle

! Vertical/bottom boundary fills (vbc*.f90, bbcw.f90); kind selects which.
      subroutine ref_vbc(kind,ni,nj,nk,a,b,c,d,e,f,g,h)                  &
     &   bind(C,name='ref_vbc')
      integer(c_int), value :: kind,ni,nj,nk
      real a(*),b(*),c(*),d(*),e(*),f(*),g(*),h(*)
      select case(kind)
      case(0)
        call s_vbcs(ni,nj,nk,a)
      case(1)
        call s_vbcu(ni,nj,nk,a)
      case(2)
        call s_vbcv(ni,nj,nk,a)
      case(3)
        call s_vbcp(idbbc,ni,nj,nk,a)
      case(4)
        call s_vbcwc(idbbc,idtbc,ni,nj,nk,a)
      case(5)
! vbcw(j31,j32,jcb8w,mf,uf,vf,wc,wf,scratch...)
        call s_vbcw(idbbc,idtbc,idmpopt,idmfcopt,ni,nj,nk,b,b,c,b,d,e,   &
     &              f,a,g,g,h)
      case(6)
! bbcw(j31,j32,mf,aa,uf,vf,wf,scratch)
        call s_bbcw(idmpopt,idmfcopt,ni,nj,nk,b,b,b,c,d,e,a,g,h)
      end select
      end subroutine ref_vbc

      subroutine ref_timeflt(ni,nj,nk,up,uf,vp,vf,wp,wf,ppp,ppf,ptpp,    &
     &   ptpf,qvp,qvf,qwtrp,qwtrf,qicep,qicef,nicep,nicef,u,v,w,pp,ptp,  &
     &   qv,qwtr,qice,nice,dmy) bind(C,name='ref_timeflt')
      integer(c_int), value :: ni,nj,nk
      real up(*),uf(*),vp(*),vf(*),wp(*),wf(*),ppp(*),ppf(*),ptpp(*)
      real ptpf(*),qvp(*),qvf(*),qwtrp(*),qwtrf(*),qicep(*),qicef(*)
      real nicep(*),nicef(*),u(*),v(*),w(*),pp(*),ptp(*),qv(*),qwtr(*)
      real qice(*),nice(*),dmy(*)
      integer nqa(0:4),land(1)
      nqa(:)=1
      land(1)=0
      call s_timeflt(idsfcopt,idcphopt,idhaiopt,idqcgopt,idaslopt,       &
     &               idtrkopt,idtubopt,idfilcoe,'moist',0.e0,ni,nj,nk,   &
     &               2,1,3,1,nqa,1,land,up,uf,vp,vf,wp,wf,ppp,ppf,       &
     &               ptpp,ptpf,qvp,qvf,qwtrp,qwtrf,dmy,dmy,qicep,qicef,  &
     &               nicep,nicef,dmy,dmy,dmy,dmy,dmy,dmy,dmy,dmy,dmy,    &
     &               dmy,dmy,dmy,u,v,w,pp,ptp,qv,qwtr,dmy,qice,nice,     &
     &               dmy,dmy,dmy,dmy,dmy,dmy)
      end subroutine ref_timeflt

      subroutine ref_totalqwi(ni,nj,nk,qwtr,qice,qall)                   &
     &   bind(C,name='ref_totalqwi')
      integer(c_int), value :: ni,nj,nk
      real qwtr(*),qice(*),qall(*)
      call s_totalqwi(idcphopt,idhaiopt,ni,nj,nk,2,3,qwtr,qice,qall)
      end subroutine ref_totalqwi


      subroutine ref_setref() bind(C,name='ref_setref')
      call s_setref
      end subroutine ref_setref

      subroutine ref_micro_prologue(ni,nj,nk,rbr,pbr,pp,qw,qi,rbv,pi,p,  &
     &   nw,nic) bind(C,name='ref_micro_prologue')
      integer(c_int), value :: ni,nj,nk
      real rbr(*),pbr(*),pp(*),qw(*),qi(*),rbv(*),pi(*),p(*),nw(*),nic(*)
      call s_getvdens(ni,nj,nk,rbr,rbv)
      call s_getexner(ni,nj,nk,pbr,pp,pi,p)
      call s_diagnw(ni,nj,nk,2,2,rbr,qw,nw)
      call s_diagnsg(idhaiopt,ni,nj,nk,3,3,rbr,rbv,qi,nic)
      end subroutine ref_micro_prologue

      subroutine ref_adjstnci(ni,nj,nk,qi,nci) bind(C,name='ref_adjstnci')
      integer(c_int), value :: ni,nj,nk
      real qi(*),nci(*)
      call s_adjstnci(ni,nj,nk,qi,nci)
      end subroutine ref_adjstnci

! termblk: q(1:6) = qc,qr,qi,qs,qg,(qh=qg); nc(1:6) likewise; u(1:12) out
      subroutine ref_termblk(ni,nj,nk,rbv,qc,qr,qi,qs,qg,ncc,ncr,nci,    &
     &   ncs,ncg,ucq,urq,uiq,usq,ugq,ucn,urn,uin,usn,ugn,dmy)             &
     &   bind(C,name='ref_termblk')
      integer(c_int), value :: ni,nj,nk
      real rbv(*),qc(*),qr(*),qi(*),qs(*),qg(*),ncc(*),ncr(*),nci(*)
      real ncs(*),ncg(*),ucq(*),urq(*),uiq(*),usq(*),ugq(*),ucn(*),urn(*)
      real uin(*),usn(*),ugn(*),dmy(*)
      call s_termblk(idcphopt,idhaiopt,idthresq,ni,nj,nk,rbv,qc,qr,qi,   &
     &               qs,qg,qg,ncc,ncr,nci,ncs,ncg,ncg,ucq,urq,uiq,usq,   &
     &               ugq,dmy,ucn,urn,uin,usn,ugn,dmy)
      end subroutine ref_termblk

! srcblk: inp = 22 fields (ptbr,rbr,rbv,pi,p,ptpp,qvp,qcp,qrp,qip,qsp,qgp,
! nccp,ncrp,ncip,ncsp,ncgp,urq,usq,ugq, dummy, dummy),
! fut = 8 fields (ptpf,qvf,qcf,qrf,qif,qsf,qgf,ncif), wrk = 80 fields.
      subroutine ref_srcblk(ni,nj,nk,dtb,thresq,inp,fut,wrk)              &
     &   bind(C,name='ref_srcblk')
      integer(c_int), value :: ni,nj,nk
      real, value :: dtb,thresq
      real inp(*),fut(*),wrk(*)
      integer n
      n=(ni+2)*(nj+2)*nk
      call s_srcblk(2,0,0,dtb,thresq,ni,nj,nk,                            &
     &  inp(1),inp(1+n),inp(1+2*n),inp(1+3*n),inp(1+4*n),inp(1+5*n),       &
     &  inp(1+6*n),inp(1+7*n),inp(1+8*n),inp(1+9*n),inp(1+10*n),           &
     &  inp(1+11*n),inp(1+11*n),                                           &
     &  inp(1+12*n),inp(1+13*n),inp(1+14*n),inp(1+15*n),inp(1+16*n),       &
     &  inp(1+16*n),inp(1+20*n),                                           &
     &  inp(1+17*n),inp(1+18*n),inp(1+19*n),inp(1+19*n),                   &
     &  inp(1+21*n),inp(1+21*n),inp(1+21*n),inp(1+21*n),                   &
     &  fut(1),fut(1+n),fut(1+2*n),fut(1+3*n),fut(1+4*n),fut(1+5*n),       &
     &  fut(1+6*n),fut(1+6*n),wrk(1+70*n),wrk(1+70*n),fut(1+7*n),          &
     &  wrk(1+71*n),wrk(1+72*n),wrk(1+73*n),                               &
     &  wrk(1+74*n),wrk(1+74*n),wrk(1+74*n),wrk(1+74*n),wrk(1+74*n),       &
     &  wrk(1+74*n),                                                       &
     &  wrk(1),wrk(1+n),wrk(1+2*n),wrk(1+3*n),wrk(1+4*n),wrk(1+5*n),       &
     &  wrk(1+6*n),wrk(1+7*n),wrk(1+8*n),wrk(1+9*n),wrk(1+10*n),           &
     &  wrk(1+11*n),wrk(1+12*n),wrk(1+13*n),wrk(1+14*n),wrk(1+15*n),       &
     &  wrk(1+16*n),wrk(1+17*n),wrk(1+18*n),wrk(1+19*n),wrk(1+20*n),       &
     &  wrk(1+21*n),wrk(1+22*n),wrk(1+23*n),wrk(1+24*n),wrk(1+25*n),       &
     &  wrk(1+26*n),wrk(1+27*n),wrk(1+28*n),wrk(1+29*n),wrk(1+30*n),       &
     &  wrk(1+31*n),wrk(1+32*n),wrk(1+33*n),wrk(1+34*n),wrk(1+35*n),       &
     &  wrk(1+36*n),wrk(1+37*n),wrk(1+38*n),wrk(1+39*n),wrk(1+40*n),       &
     &  wrk(1+41*n),wrk(1+42*n),wrk(1+43*n),wrk(1+44*n),wrk(1+45*n),       &
     &  wrk(1+46*n),wrk(1+47*n),wrk(1+48*n),wrk(1+49*n),wrk(1+50*n),       &
     &  wrk(1+51*n),wrk(1+52*n),wrk(1+53*n),wrk(1+54*n),wrk(1+55*n),       &
     &  wrk(1+56*n),wrk(1+57*n),wrk(1+58*n),wrk(1+59*n),wrk(1+60*n))
      end subroutine ref_srcblk

      subroutine ref_upwqp(dtp,ni,nj,nk,rbr,rst,uq,qpf,precip,flx)        &
     &   bind(C,name='ref_upwqp')
      real, value :: dtp
      integer(c_int), value :: ni,nj,nk
      real rbr(*),rst(*),uq(*),qpf(*),precip(*),flx(*)
      call s_upwqp(idadvopt,iddziv,dtp,ni,nj,nk,rbr,rst,uq,qpf,precip,   &
     &             flx)
      end subroutine ref_upwqp

      subroutine ref_upwnp(dtp,ni,nj,nk,rbr,rst,un,ncf,flx)               &
     &   bind(C,name='ref_upwnp')
      real, value :: dtp
      integer(c_int), value :: ni,nj,nk
      real rbr(*),rst(*),un(*),ncf(*),flx(*)
      call s_upwnp(iddziv,dtp,ni,nj,nk,rbr,rst,un,ncf,flx)
      end subroutine ref_upwnp

! satadjst: q(1:6) = qv,qc,qr,qi,qs,qg stacked; nice(1:3); nwtr(1:2)
      subroutine ref_satadjst(ni,nj,nk,pbr,ptbr,w,pp,ptp,q,nw,nic,pi,p)  &
     &   bind(C,name='ref_satadjst')
      integer(c_int), value :: ni,nj,nk
      real pbr(*),ptbr(*),w(*),pp(*),ptp(*),q(*),nw(*),nic(*),pi(*),p(*)
      integer n
      n=(ni+2)*(nj+2)*nk
      call s_satadjst(idcphopt,ni,nj,nk,2,2,3,3,pbr,ptbr,w,pp,ptp,q(1),  &
     &                q(1+n),nw,q(1+3*n),nic,pi,p)
      end subroutine ref_satadjst

! Open lateral boundaries (wbc..nbc >= 4): phase speeds and radiation
! fills with the namelist values ref_configure left (exbopt=0, nggopt=0,
! lspopt=0, lbcvar/gpvvar all 'x', lbnews=lbnorm=0).  gwave is set here.
      subroutine ref_phvuvw(gwave,dtb,dts,ni,nj,nk,rmf,rmf8u,rmf8v,      &
     &   u,ucpx,ucpy,vcpx,vcpy,wcpx,wcpy,u8v,v8u,cpavex,cpavey)          &
     &   bind(C,name='ref_phvuvw')
      real, value :: gwave,dtb,dts
      integer(c_int), value :: ni,nj,nk
      real rmf(*),rmf8u(*),rmf8v(*),u(*)
      real ucpx(*),ucpy(*),vcpx(*),vcpy(*),wcpx(*),wcpy(*)
      real u8v(*),v8u(*),cpavex(*),cpavey(*)
      rname(idgwave)=gwave
      call s_phvuvw(idexbvar,idexbopt,idwbc,idebc,idsbc,idnbc,idmpopt,   &
     &              idmfcopt,iddxiv,iddyiv,idgwave,dtb,dts,ni,nj,nk,     &
     &              rmf,rmf8u,rmf8v,u,u,u,u,u,u,u,u,u,                   &
     &              ucpx,ucpy,vcpx,vcpy,wcpx,wcpy,u8v,v8u,cpavex,cpavey)
      end subroutine ref_phvuvw

      subroutine ref_phvs(gwave,sml,dtb,dts,ni,nj,nk,rmf,u,scpx,scpy,    &
     &   cpavex,cpavey) bind(C,name='ref_phvs')
      real, value :: gwave,dtb,dts
      integer(c_int), value :: sml,ni,nj,nk
      real rmf(*),u(*),scpx(*),scpy(*),cpavex(*),cpavey(*)
      character(len=3) fproc
      rname(idgwave)=gwave
      fproc='big'
      if(sml.ne.0) fproc='sml'
      call s_phvs(idexbvar,idexbopt,idwbc,idebc,idsbc,idnbc,idadvopt,    &
     &            idmpopt,idmfcopt,iddxiv,iddyiv,idgwave,fproc,4,dtb,dts, &
     &            dts,ni,nj,nk,rmf,u,u,u,u,u,scpx,scpy,cpavex,cpavey)
      end subroutine ref_phvs

! kind: 1 rbcu, 2 rbcv, 3 rbcw, 4 rbcss (one field updated in place, its
! base/forcing companions dummies)
      subroutine ref_rbcuvw(kind,dts,ni,nj,nk,cpx,cpy,dum,var)           &
     &   bind(C,name='ref_rbcuvw')
      integer(c_int), value :: kind,ni,nj,nk
      real, value :: dts
      real cpx(*),cpy(*),dum(*),var(*)
      if(kind.eq.1) then
        call s_rbcu(idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt,idlspopt,  &
     &              idvspopt,idlbnews,idlbnorm,1,dts,0.e0,ni,nj,nk,dum,  &
     &              cpx,cpy,dum,dum,var)
      else if(kind.eq.2) then
        call s_rbcv(idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt,idlspopt,  &
     &              idvspopt,idlbnews,idlbnorm,1,dts,0.e0,ni,nj,nk,dum,  &
     &              cpx,cpy,dum,dum,var)
      else if(kind.eq.3) then
        call s_rbcw(idgpvvar,idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt,  &
     &              idlspopt,idvspopt,idlbnews,1,dts,0.e0,ni,nj,nk,      &
     &              cpx,cpy,dum,dum,var)
      else
        call s_rbcss(idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt,idlspopt, &
     &               idvspopt,idlbnews,4,1,dts,0.e0,ni,nj,nk,cpx,cpy,    &
     &               dum,dum,var)
      end if
      end subroutine ref_rbcuvw

! kind: 1 rbcpt, 2 rbcqv, 3 rbcq (qc slot), 4 rbcs (nci slot)
      subroutine ref_rbclf(kind,dt,ni,nj,nk,s,sp,cpx,cpy,dum,sf)         &
     &   bind(C,name='ref_rbclf')
      integer(c_int), value :: kind,ni,nj,nk
      real, value :: dt
      real s(*),sp(*),cpx(*),cpy(*),dum(*),sf(*)
      if(kind.eq.1) then
        call s_rbcpt(idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt,idlspopt, &
     &               idvspopt,idadvopt,idlbnews,1,dt,0.e0,ni,nj,nk,s,sp, &
     &               cpx,cpy,dum,dum,sf)
      else if(kind.eq.2) then
        call s_rbcqv(idgpvvar,idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt, &
     &               idlspopt,idvspopt,idadvopt,idlbnews,1,dt,0.e0,      &
     &               ni,nj,nk,dum,s,sp,cpx,cpy,dum,dum,sf)
      else if(kind.eq.3) then
        call s_rbcq(idgpvvar,idlbcvar,idwbc,idebc,idsbc,idnbc,idnggopt,  &
     &              idlspopt,idvspopt,idadvopt,idlbnews,3,7,1,dt,0.e0,   &
     &              ni,nj,nk,s,sp,cpx,cpy,dum,dum,sf)
      else
        call s_rbcs(idlbcvar,idwbc,idebc,idsbc,idnbc,idadvopt,idlbnews,  &
     &              7,dt,ni,nj,nk,s,sp,cpx,cpy,sf)
      end if
      end subroutine ref_rbclf

      subroutine ref_bc4news(isw,ise,jss,jsn,ni,nj,kmax,var)             &
     &   bind(C,name='ref_bc4news')
      integer(c_int), value :: isw,ise,jss,jsn,ni,nj,kmax
      real var(*)
      call s_bc4news(idwbc,idebc,idsbc,idnbc,isw,ise,jss,jsn,ni,nj,kmax, &
     &               var)
      end subroutine ref_bc4news

! kind: 0 bc8u, 1 bc8v, 2 bc8w (kmax = nk)
      subroutine ref_bc8(kind,ni,nj,kmax,var) bind(C,name='ref_bc8')
      integer(c_int), value :: kind,ni,nj,kmax
      real var(*)
      select case(kind)
      case(0)
        call s_bc8u(idwbc,idebc,ni,nj,kmax,var)
      case(1)
        call s_bc8v(idsbc,idnbc,ni,nj,kmax,var)
      case default
        call s_bc8w(idbbc,idtbc,ni,nj,kmax,var)
      end select
      end subroutine ref_bc8

! kind: 1 lbcu, 2 lbcv, 3 lbcw, 4 lbcs, 5 lbcwc
      subroutine ref_lbc(kind,ni,nj,nk,var) bind(C,name='ref_lbc')
      integer(c_int), value :: kind,ni,nj,nk
      real var(*)
      select case(kind)
      case(1)
        call s_lbcu(idwbc,idebc,idsbc,idnbc,idadvopt,ni,nj,nk,var)
      case(2)
        call s_lbcv(idwbc,idebc,idsbc,idnbc,idadvopt,ni,nj,nk,var)
      case(3)
        call s_lbcw(idwbc,idebc,idsbc,idnbc,idadvopt,ni,nj,nk,var)
      case(4)
        call s_lbcs(idwbc,idebc,idsbc,idnbc,idadvopt,ni,nj,nk,var)
      case default
        call s_lbcwc(idwbc,idebc,idsbc,idnbc,ni,nj,nk,var)
      end select
      end subroutine ref_lbc

! 4th-order smoothing (smtopt 2/3) with the coefficients set here.
! kind: 0 smoo4s, 1 smoo4qv, 2 smoo4p, 3 smoo4uvw (u,v,w in a..c, their
! forcings in f1..f3, base/metrics in m1..m8)
      subroutine ref_smoo4(kind,smtopt,smh,smv,ni,nj,nk,a,b,c,f1,f2,f3,   &
     &   m1,m2,m3,m4,m5,m6,m7,m8,t1,t2,t3,t4,t5) bind(C,name='ref_smoo4')
      integer(c_int), value :: kind,smtopt,ni,nj,nk
      real, value :: smh,smv
      real a(*),b(*),c(*),f1(*),f2(*),f3(*)
      real m1(*),m2(*),m3(*),m4(*),m5(*),m6(*),m7(*),m8(*)
      real t1(*),t2(*),t3(*),t4(*),t5(*)
      iname(idsmtopt)=smtopt
      rname(idsmhcoe)=smh
      rname(idsmvcoe)=smv
      select case(kind)
      case(0)
        call s_smoo4s(idsmtopt,idiwest,idieast,idjsouth,idjnorth,        &
     &                idsmhcoe,idsmvcoe,ni,nj,nk,m1,a,f1,t4,t5,t1,t2,t3)
      case(1)
        call s_smoo4qv(idsmtopt,idiwest,idieast,idjsouth,idjnorth,       &
     &                 idsmhcoe,idsmvcoe,ni,nj,nk,m2,m1,a,f1,t4,t5,      &
     &                 t1,t2,t3)
      case(2)
        call s_smoo4p(idsmtopt,idiwest,idieast,idjsouth,idjnorth,        &
     &                idsmhcoe,idsmvcoe,ni,nj,nk,a,f1,t5,t1,t2,t3)
      case default
        call s_smoo4uvw(idsmtopt,idiwest,idieast,idjsouth,idjnorth,      &
     &                  idsmhcoe,idsmvcoe,ni,nj,nk,m1,m2,m3,m4,m5,m6,m7,  &
     &                  m8,a,b,c,f1,f2,f3,t1,t2,t3,t4,t5)
      end select
      end subroutine ref_smoo4

! The first four calls of s_turbdrv (turbdrv.f90:245-276), tubopt=2,
! isoopt=1, with separate work arrays (w1..w4 3-D, w5 2-D).
      subroutine ref_turbcoef(moist,dtb,ni,nj,nk,jcb,jcb8u,jcb8v,jcb8w,  &
     &   pbr,ptbr,rbr,ppp,ptpp,qvp,qallp,up,vp,wp,tkep,t11,t22,t33,t12,   &
     &   t13,t23,t31,t32,ssq,nsq8w,rkh,rkv,priv,w1,w2,w3,w4,w5,zph,mf)    &
     &   bind(C,name='ref_turbcoef')
      integer(c_int), value :: moist,ni,nj,nk
      real, value :: dtb
      real jcb(*),jcb8u(*),jcb8v(*),jcb8w(*),pbr(*),ptbr(*),rbr(*)
      real ppp(*),ptpp(*),qvp(*),qallp(*),up(*),vp(*),wp(*),tkep(*)
      real t11(*),t22(*),t33(*),t12(*),t13(*),t23(*),t31(*),t32(*)
      real ssq(*),nsq8w(*),rkh(*),rkv(*),priv(*)
      real w1(*),w2(*),w3(*),w4(*),w5(*),zph(*),mf(*)
      character(len=5) fmois
      iname(idtubopt)=2
      iname(idisoopt)=1
      fmois='dry  '
      if(moist.ne.0) fmois='moist'
      call s_bruntv(idcphopt,iddziv,idthresq,fmois,ni,nj,nk,jcb8w,       &
     &              pbr,ptbr,ppp,ptpp,qvp,qallp,nsq8w,w1,w2,w3,w5)
      call s_defomten(idtrnopt,idmpopt,idmfcopt,idiwest,idieast,         &
     &                idjsouth,idjnorth,iddxiv,iddyiv,iddziv,ni,nj,nk,   &
     &                zph,zph,jcb,jcb8u,jcb8v,jcb8w,mf,mf,mf,up,vp,wp,   &
     &                t11,t22,t33,t12,t13,t23,t31,t32,w1,w2,w3,w4)
      call s_defomssq(ni,nj,nk,t11,t22,t33,t12,t31,t32,ssq)
      call s_eddyvis(idmpopt,idmfcopt,idsfcopt,idtubopt,idisoopt,        &
     &               iddx,iddy,iddz,dtb,ni,nj,nk,zph,jcb,mf,rbr,         &
     &               tkep,ssq,nsq8w,rkh,rkv,priv)
      iname(idtubopt)=0
      end subroutine ref_turbcoef

! s_turbdrv itself (turbdrv.f90:80-85), tubopt=2, isoopt=1.
      subroutine ref_turbdrv(moist,dtb,ni,nj,nk,jcb,jcb8u,jcb8v,jcb8w,   &
     &   pbr,ptbr,rbr,up,vp,wp,ppp,ptpp,qvp,tkep,qallp,ufrc,vfrc,wfrc,    &
     &   rkh,rkv,priv,ssq,nsq8w,t13,t23,t33,t11,t22,t12,t31,t32,w1,zph,   &
     &   mf) bind(C,name='ref_turbdrv')
      integer(c_int), value :: moist,ni,nj,nk
      real, value :: dtb
      real jcb(*),jcb8u(*),jcb8v(*),jcb8w(*),pbr(*),ptbr(*),rbr(*)
      real up(*),vp(*),wp(*),ppp(*),ptpp(*),qvp(*),tkep(*),qallp(*)
      real ufrc(*),vfrc(*),wfrc(*),rkh(*),rkv(*),priv(*),ssq(*),nsq8w(*)
      real t13(*),t23(*),t33(*),t11(*),t22(*),t12(*),t31(*),t32(*)
      real w1(*),zph(*),mf(*)
      character(len=5) fmois
      iname(idtubopt)=2
      iname(idisoopt)=1
      fmois='dry  '
      if(moist.ne.0) fmois='moist'
      call s_turbdrv(fmois,dtb,ni,nj,nk,zph,zph,zph,jcb,jcb8u,jcb8v,     &
     &               jcb8w,mf,mf,mf,mf,mf,mf,pbr,ptbr,rbr,up,vp,wp,ppp,  &
     &               ptpp,qvp,tkep,qallp,ufrc,vfrc,wfrc,rkh,rkv,priv,    &
     &               ssq,nsq8w,t13,t23,t33,t11,t22,t12,t31,t32,w1)
      iname(idtubopt)=0
      end subroutine ref_turbdrv

      subroutine ref_turbflx(ni,nj,nk,jcb,s,rkh8u,rkh8v,rkv8w,h1,h2,h3,  &
     &   w1,w2,w3,w4) bind(C,name='ref_turbflx')
      integer(c_int), value :: ni,nj,nk
      real jcb(*),s(*),rkh8u(*),rkh8v(*),rkv8w(*),h1(*),h2(*),h3(*)
      real w1(*),w2(*),w3(*),w4(*)
      call s_turbflx(idtrnopt,idsfcopt,iddxiv,iddyiv,iddziv,ni,nj,nk,    &
     &               w1,w1,jcb,s,w2,rkh8u,rkh8v,rkv8w,h1,h2,h3,w3,w3,w4)
      end subroutine ref_turbflx

      subroutine ref_turbs(ni,nj,nk,jcb8u,jcb8v,h1,h2,h3,sfrc,w1,w2,w3, &
     &   w4) bind(C,name='ref_turbs')
      integer(c_int), value :: ni,nj,nk
      real jcb8u(*),jcb8v(*),h1(*),h2(*),h3(*),sfrc(*)
      real w1(*),w2(*),w3(*),w4(*)
      call s_turbs(idtrnopt,idmpopt,idmfcopt,iddxiv,iddyiv,iddziv,       &
     &             ni,nj,nk,w1,w1,jcb8u,jcb8v,w4,w4,w4,w4,h1,h2,h3,sfrc, &
     &             w1,w2,w3)
      end subroutine ref_turbs

      subroutine ref_kh8uv(ni,nj,nk,rkh,rkh8u,rkh8v,w1)                 &
     &   bind(C,name='ref_kh8uv')
      integer(c_int), value :: ni,nj,nk
      real rkh(*),rkh8u(*),rkh8v(*),w1(*)
      call s_kh8uv(idmpopt,idmfcopt,ni,nj,nk,w1,rkh,rkh8u,rkh8v)
      end subroutine ref_kh8uv

! s_forcetke (forcetke.f90:88-94), tubopt=2, isoopt=1, with the lateral
! sponge and smoothing off (ref_configure); j31, j32 and the map factors
! are not read for trnopt=0, mfcopt=0 and get the work array w1.
      subroutine ref_forcetke(ksp02,ni,nj,nk,jcb,jcb8u,jcb8v,jcb8w,rbr,  &
     &   rst,rstxu,rstxv,rstxwc,tke,tkep,rbct,priv,nsq8w,ssq,rkh,rkv,    &
     &   tkefrc,h3,rkv8s,w1,w2,w3,rbcxy) bind(C,name='ref_forcetke')
      integer(c_int), value :: ksp02,ni,nj,nk
      real jcb(*),jcb8u(*),jcb8v(*),jcb8w(*),rbr(*),rst(*),rstxu(*)
      real rstxv(*),rstxwc(*),tke(*),tkep(*),rbct(*),priv(*),nsq8w(*)
      real ssq(*),rkh(*),rkv(*),tkefrc(*),h3(*),rkv8s(*),w1(*),w2(*)
      real w3(*),rbcxy(*)
      integer ksp0(2)
      ksp0(1)=2
      ksp0(2)=ksp02
      iname(idtubopt)=2
      iname(idisoopt)=1
      call s_forcetke(idlspvar,idvspvar,idlspopt,idvspopt,idsmtopt,      &
     &                ksp0,ni,nj,nk,w1,w1,jcb,jcb8u,jcb8v,jcb8w,w1,w1,   &
     &                w1,w1,rbr,rst,rstxu,rstxv,rstxwc,tke,tkep,rbcxy,   &
     &                rbct,priv,nsq8w,ssq,rkh,rkv,tkefrc,h3,rkv8s,w1,w2, &
     &                w3)
      iname(idtubopt)=0
      end subroutine ref_forcetke

      end module cress_ref_harness
