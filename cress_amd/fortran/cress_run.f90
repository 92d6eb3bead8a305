!-----------------------------------------------------------------------
! cress_run -- Fortran 90 host of the MI355X large step.
!
!   cress_run <user.conf> <state.bin> <out.bin> [--check]
!
! With --check it prints the resolved options and stops before touching
! the GPU.
!
! Reads the run options from the reference's user.conf namelist groups
! (group and member names as declared in SrcOrig/defname.f90:806-862), the
! base state and initial levels from state.bin, advances the model
! nint((etime-stime)/dtbig) large steps on the GPU through the C ABI, and
! writes the present level of every prognostic and the precipitation
! arrays to out.bin.
!
! state.bin / out.bin (stream, native endianness): a header of 8 default
! integers  'CRSS' tag(1), version(1), ni, nj, nk, ksp0(2), nrec
! then nrec records, each  field id, level, nfloat (3 integers) followed
! by nfloat default reals, in the field ids of m_cress_hip.
!
! Option handling follows the reference: on one rank the periodic lateral
! flags 1 become -1 (chknlslv.f90:751-790), the acoustic substep count is
! nsstp = 20*int(1.e2*(dtbig+.001))/int(1.e3*(dtsml+.0001)) for advopt<=3
! (slvstep.f90:174-176), and the dimensions are ni=xdim, nj=ydim, nk=zdim
! (setdim.f90:196-199 with one sub-domain).  Fatal conditions stop with a
! message, as the reference's destroy('stop') does.
!-----------------------------------------------------------------------
program cress_run

  use iso_c_binding
  use m_cress_hip

  implicit none

! namelist members (SrcOrig/defname.f90)
  integer numpe, xgroup, ygroup, xsub, ysub, xdim, ydim, zdim
  real dx, dy, dz, ulat, ulon, riu, rju
  real zsfc, zflat, dzmin, layer1, layer2
  integer sthopt
  integer trnopt
  real mnthgh(1:2), mntwx, mntwy, mntcx, mntcy
  character(len=22) sfcast
  real stime, etime
  integer wbc, ebc, sbc, nbc, bbc, tbc
  character(len=108) lbcvar
  real lbnews, lbnorm, gwave
  character(len=108) gpvvar, nggvar, exbvar, lspvar, vspvar
  real gpvitv, aslitv, gsmcoe, nggcoe, nggdlt, nggstr, nggend, nggc20
  real exnews, exnorm, lspsmt, lsnews, lsnorm, vspgpv, vspbar, botgpv
  real botbar
  integer gsmopt, gsmcnt, nggopt, exbopt, exbwid, lspopt, wdnews
  integer wdnorm, vspopt
  real dtbig, dtsml, gsdeps, weicoe, filcoe, dtvcul
  integer gwmopt, impopt, advopt
  integer smtopt
  real smhcoe, smvcoe, nlhcoe, nlvcoe
  integer mfcopt, coropt, buyopt, divopt
  integer cphopt, haiopt, ncbinw, qcgopt
  real thresq, dtcmph(1:6), bbinw, sbinw, eledlt
  integer tubopt, isoopt

  namelist /dimset/ numpe,xgroup,ygroup,xsub,ysub,xdim,ydim,zdim
  namelist /gridset/ dx,dy,dz,ulat,ulon,riu,rju
  namelist /gridsth/ zsfc,zflat,sthopt,dzmin,layer1,layer2
  namelist /terrain/ trnopt,mnthgh,mntwx,mntwy,mntcx,mntcy
  namelist /flength/ sfcast,stime,etime
  namelist /boundry/ wbc,ebc,sbc,nbc,bbc,tbc,lbcvar,lbnews,lbnorm,gwave
  namelist /gpvpram/ gpvvar,gpvitv,aslitv,gsmopt,gsmcnt,gsmcoe,nggopt,  &
     &                   nggvar,nggcoe,nggdlt,nggstr,nggend,nggc20,       &
     &                   exbopt,exbvar,exnews,exnorm,exbwid,lspopt,       &
     &                   lspvar,lspsmt,lsnews,lsnorm,wdnews,wdnorm,       &
     &                   vspopt,vspvar,vspgpv,vspbar,botgpv,botbar
  namelist /integrat/ dtbig,dtsml,gwmopt,impopt,advopt,gsdeps,weicoe,   &
     &                    filcoe,dtvcul
  namelist /smoother/ smtopt,smhcoe,smvcoe,nlhcoe,nlvcoe
  namelist /mapfcter/ mfcopt
  namelist /coriolis/ coropt
  namelist /buoyancy/ buyopt
  namelist /ddamping/ divopt
  namelist /cloudphy/ cphopt,haiopt,thresq,dtcmph,ncbinw,bbinw,sbinw,   &
     &                    qcgopt,eledlt
  namelist /turbulen/ tubopt,isoopt

  type(cress_config) :: cfg
  type(c_ptr) :: ctx
  character(len=512) :: confn, stfn, outfn, arg4
  integer :: hdr(8), rec(3)
  integer :: ios, n, nrec, nsteps, first
  integer(c_int) :: rc
  real(c_float), allocatable :: buf(:)
  integer, parameter :: nprog = 12
  integer(c_int), parameter :: prog(nprog) = (/ fid_u, fid_v, fid_w,    &
     &  fid_pp, fid_ptp, fid_qv, fid_qc, fid_qr, fid_qi, fid_qs, fid_qg, &
     &  fid_nci /)
  integer(c_size_t) :: nf

  if(command_argument_count() < 3 .or. command_argument_count() > 4) then
    write(*,'(a)') 'usage: cress_run <user.conf> <state.bin> <out.bin> [--check]'
    stop 2
  end if
  call get_command_argument(1, confn)
  call get_command_argument(2, stfn)
  call get_command_argument(3, outfn)
  arg4 = ' '
  if(command_argument_count() == 4) call get_command_argument(4, arg4)

! defaults of the options this host does not require in user.conf
  numpe = 1; xgroup = 1; ygroup = 1; xsub = 1; ysub = 1
  sthopt = 0; trnopt = 0; stime = 0.0; mfcopt = 0; coropt = 0
  smtopt = 0; tubopt = 0; buyopt = 1; divopt = 0; haiopt = 0
  gwmopt = 0; impopt = 1; weicoe = 0.6; filcoe = 0.1
  vspopt = 0; vspvar = 'xxxxxxxxxx'; thresq = 1.e-8
  lbcvar = 'xxxxxxxxxx'; lbnews = 0.0; lbnorm = 0.0; gwave = 0.0

  open(10, file=trim(confn), status='old', action='read', iostat=ios)
  if(ios /= 0) call fatal('cannot open '//trim(confn))
  read(10, nml=dimset, iostat=ios)
  if(ios /= 0) call fatal('namelist dimset')
  call rd_opt('gridset')
  call rd_opt('gridsth')
  call rd_opt('terrain')
  call rd_opt('flength')
  call rd_opt('boundry')
  call rd_opt('gpvpram')
  call rd_opt('integrat')
  call rd_opt('smoother')
  call rd_opt('mapfcter')
  call rd_opt('coriolis')
  call rd_opt('buoyancy')
  call rd_opt('ddamping')
  call rd_opt('cloudphy')
  call rd_opt('turbulen')
  close(10)

  if(numpe /= 1 .or. xsub*xgroup /= 1 .or. ysub*ygroup /= 1) then
    call fatal('one GPU runs the whole domain: numpe, xsub, ysub, '//   &
     &         'xgroup, ygroup must be 1')
  end if
  if(trnopt /= 0 .or. mfcopt /= 0 .or. tubopt /= 0) then
    call fatal('trnopt, mfcopt and tubopt must be 0')
  end if
  if(smtopt /= 0 .and. smtopt /= 2 .and. smtopt /= 3) then
    call fatal('smtopt must be 0, 2 or 3')
  end if

! open sides: the constant-phase-speed radiation (option 7) without the
! lateral damping toward the base state (lbcvar 'o' with lbnews/lbnorm)
  if(index(lbcvar(1:10), 'o') > 0 .and. (lbnews /= 0.0 .or. lbnorm /= 0.0)) then
    call fatal('lateral boundary damping (lbcvar/lbnews/lbnorm) is not supported')
  end if

! one rank: periodic flags become -1 (chknlslv.f90:751-790)
  if(wbc == 1) wbc = -1
  if(ebc == 1) ebc = -1
  if(sbc == 1) sbc = -1
  if(nbc == 1) nbc = -1

! state header
  open(11, file=trim(stfn), access='stream', form='unformatted',        &
     &     status='old', action='read', iostat=ios)
  if(ios /= 0) call fatal('cannot open '//trim(stfn))
  read(11, iostat=ios) hdr
  if(ios /= 0 .or. hdr(1) /= 1397969475 .or. hdr(2) /= 1)               &
     &  call fatal('not a state file: '//trim(stfn))
  if(hdr(3) /= xdim .or. hdr(4) /= ydim .or. hdr(5) /= zdim)            &
     &  call fatal('state dimensions differ from dimset')

  cfg%ni = xdim
  cfg%nj = ydim
  cfg%nk = zdim
  cfg%wbc = wbc
  cfg%ebc = ebc
  cfg%sbc = sbc
  cfg%nbc = nbc
  cfg%bbc = bbc
  cfg%tbc = tbc
  cfg%advopt = advopt
  cfg%impopt = impopt
  cfg%gwmopt = gwmopt
  cfg%buyopt = buyopt
  cfg%divopt = divopt
  cfg%cphopt = cphopt
  cfg%haiopt = haiopt
  cfg%vspopt = vspopt
  cfg%coropt = coropt
  cfg%sthopt = sthopt
  cfg%ksp0(1) = hdr(6)
  cfg%ksp0(2) = hdr(7)
  do n = 1, 10
    cfg%vspvar(n) = vspvar(n:n)
  end do
  if(advopt <= 3) then
    cfg%nsstp = 20*int(1.e2*(dtbig+.001e0))/int(1.e3*(dtsml+.0001e0))
  else
    call fatal('advopt must be 1, 2 or 3')
  end if
  cfg%dx = dx
  cfg%dy = dy
  cfg%dz = dz
  cfg%dtbig = dtbig
  cfg%dtsml = dtsml
  cfg%weicoe = weicoe
  cfg%filcoe = filcoe
  cfg%thresq = thresq
  cfg%gwave = gwave
  cfg%smtopt = smtopt
  cfg%smhcoe = smhcoe
  cfg%smvcoe = smvcoe

  if(trim(arg4) == '--check') then
    write(*,'(a,3i8)') 'dims', cfg%ni, cfg%nj, cfg%nk
    write(*,'(a,6i4)') 'bcs', cfg%wbc, cfg%ebc, cfg%sbc, cfg%nbc,       &
     &  cfg%bbc, cfg%tbc
    write(*,'(a,10i4)') 'opts', cfg%advopt, cfg%impopt, cfg%gwmopt,     &
     &  cfg%buyopt, cfg%divopt, cfg%cphopt, cfg%haiopt, cfg%vspopt,      &
     &  cfg%coropt, cfg%sthopt
    write(*,'(a,3i6)') 'ksp0_nsstp', cfg%ksp0(1), cfg%ksp0(2), cfg%nsstp
    write(*,'(2a)') 'vspvar ', vspvar(1:10)
    write(*,'(a,11es24.16)') 'reals', cfg%dx, cfg%dy, cfg%dz, cfg%dtbig,&
     &  cfg%dtsml, cfg%weicoe, cfg%filcoe, cfg%thresq, cfg%gwave,         &
     &  cfg%smhcoe, cfg%smvcoe
    write(*,'(a,i4)') 'smtopt', cfg%smtopt
    write(*,'(a,i8)') 'nsteps', nint((etime - stime) / dtbig)
    stop
  end if

  rc = cress_ctx_create(cfg, ctx)
  call check(rc, 'cress_ctx_create')

  nrec = hdr(8)
  do n = 1, nrec
    read(11, iostat=ios) rec
    if(ios /= 0) call fatal('truncated state file')
    allocate(buf(rec(3)))
    read(11, iostat=ios) buf
    if(ios /= 0) call fatal('truncated state file')
    rc = cress_field_size(ctx, rec(1), nf)
    call check(rc, 'cress_field_size')
    if(nf /= rec(3)) call fatal('record size differs from the field')
    rc = cress_upload(ctx, rec(1), rec(2), buf)
    call check(rc, 'cress_upload')
    deallocate(buf)
  end do
  close(11)

  nsteps = nint((etime - stime) / dtbig)
  first = 0
  if(stime == 0.0) first = 1
  if(nsteps > 0) then
    rc = cress_large_step(ctx, 1, first)
    call check(rc, 'cress_large_step')
  end if
  if(nsteps > 1) then
    rc = cress_large_step(ctx, nsteps - 1, 0)
    call check(rc, 'cress_large_step')
  end if
  rc = cress_sync(ctx)
  call check(rc, 'cress_sync')

  open(12, file=trim(outfn), access='stream', form='unformatted',       &
     &     status='replace', action='write', iostat=ios)
  if(ios /= 0) call fatal('cannot create '//trim(outfn))
  n = nprog
  if(cphopt /= 2) n = 6
  hdr(8) = n
  if(cphopt == 2) hdr(8) = n + 2
  write(12) hdr
  do n = 1, nprog
    if(cphopt /= 2 .and. n > 6) exit
    call put(prog(n), 1)
  end do
  if(cphopt == 2) then
    call put(fid_prwtr, 0)
    call put(fid_price, 0)
  end if
  close(12)

  rc = cress_ctx_destroy(ctx)
  write(*,'(a,i6,a,i4,a,i5,a,i5,a,i4)') 'cress_run: ', nsteps,          &
     &  ' large steps, nsstp=', cfg%nsstp, ', grid ', xdim, ' x', ydim,  &
     &  ' x', zdim

contains

! Reads one optional namelist group; a missing group keeps the defaults.
  subroutine rd_opt(grp)
    character(len=*), intent(in) :: grp
    integer :: ios2
    rewind(10)
    select case(grp)
    case('gridset')
      read(10, nml=gridset, iostat=ios2)
    case('gridsth')
      read(10, nml=gridsth, iostat=ios2)
    case('terrain')
      read(10, nml=terrain, iostat=ios2)
    case('flength')
      read(10, nml=flength, iostat=ios2)
    case('boundry')
      read(10, nml=boundry, iostat=ios2)
    case('gpvpram')
      read(10, nml=gpvpram, iostat=ios2)
    case('integrat')
      read(10, nml=integrat, iostat=ios2)
    case('smoother')
      read(10, nml=smoother, iostat=ios2)
    case('mapfcter')
      read(10, nml=mapfcter, iostat=ios2)
    case('coriolis')
      read(10, nml=coriolis, iostat=ios2)
    case('buoyancy')
      read(10, nml=buoyancy, iostat=ios2)
    case('ddamping')
      read(10, nml=ddamping, iostat=ios2)
    case('cloudphy')
      read(10, nml=cloudphy, iostat=ios2)
    case('turbulen')
      read(10, nml=turbulen, iostat=ios2)
    end select
    if(ios2 > 0) call fatal('namelist '//grp)
  end subroutine rd_opt

  subroutine put(fid, lev)
    integer(c_int), intent(in) :: fid
    integer, intent(in) :: lev
    integer(c_int) :: rc2
    integer(c_size_t) :: m
    real(c_float), allocatable :: b(:)
    rc2 = cress_field_size(ctx, fid, m)
    call check(rc2, 'cress_field_size')
    allocate(b(m))
    rc2 = cress_download(ctx, fid, lev, b)
    call check(rc2, 'cress_download')
    write(12) int(fid), lev, int(m)
    write(12) b
    deallocate(b)
  end subroutine put

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if(code /= cress_ok) then
      call fatal(what//': '//trim(cress_errmsg(code)))
    end if
  end subroutine check

  subroutine fatal(msg)
    character(len=*), intent(in) :: msg
    write(*,'(2a)') 'cress_run: ', msg
    stop 1
  end subroutine fatal

end program cress_run
