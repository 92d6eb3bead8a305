!-----------------------------------------------------------------------
! cress_run -- Fortran 90 host of the MI355X large step.
!
!   cress_run <user.conf> <state.bin> <out.bin> [--check]
!
! With --check it prints the resolved options and stops before touching
! the GPU.
!
! With an &outfomat group in user.conf it also writes the reference's dump
! files (dmpvar, every dmpitv seconds and at etime) and, with resopt=1, its
! restart files (every resitv seconds and at etime), as m_cress_files
! describes; iniopt=2 or 12 in &initype starts from the restart file of
! time stime after the state file is loaded (the state file still supplies
! the metrics, which a restart file does not hold).  The step count follows
! slvstep.f90:168-171 and gettime.f90:122-123 (times in ms).
!
! Reads the run options from the reference's user.conf namelist groups
! (group and member names as declared in SrcOrig/defname.f90:806-862), the
! base state and initial levels from state.bin, advances the model from
! stime to etime on the GPU through the C ABI, and
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
  use m_cress_files

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
  character(len=108) exprim, crsdir, datdir
  integer iniopt, snddim, masopt
  character(len=108) sndtyp
  real zsnd0, psnd0, maseps, alpha1, alpha2
  integer dmpfmt, dmplev, dmpmon, resopt, mxnopt
  character(len=108) dmpvar, mxnvar
  real dmpitv, monitv, resitv, mxnitv

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
  namelist /runame/ exprim
  namelist /drname/ crsdir,datdir
  namelist /initype/ iniopt,snddim,sndtyp,zsnd0,psnd0,masopt,maseps,     &
     &                   alpha1,alpha2
  namelist /outfomat/ dmpfmt,dmplev,dmpmon,dmpvar,dmpitv,monitv,         &
     &                    resopt,resitv,mxnopt,mxnvar,mxnitv

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
  logical :: outf, moist, opsides, lastfirst
  integer(8) :: dtms, ctime, ftime, nb0, nb1, ib, etms
  real(c_float), allocatable :: qvb(:,:,:)

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
  exprim = 'cress'; crsdir = './'; iniopt = 1
  dmpfmt = 2; dmplev = 1; dmpmon = 0; dmpvar = repeat('x', 108)
  dmpitv = 0.0; monitv = 0.0; resopt = 0; resitv = 0.0

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
  call rd_opt('runame')
  call rd_opt('drname')
  call rd_opt('initype')
  outf = .false.
  call rd_opt('outfomat')
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

  if(outf) then
    if(dmpfmt /= 2 .or. dmplev /= 1 .or. dmpmon /= 0) then
      call fatal('dumps: only dmpfmt=2, dmplev=1, dmpmon=0 are written')
    end if
    if(dmpvar(15:15) == 'o') call fatal('dumps: dmpvar(15:15) (zph) is not held')
    if(dmpitv < 1.0 .or. (resopt == 1 .and. resitv < 1.0)) then
      call fatal('dmpitv and resitv must be at least 1 s')
    end if
  end if
  if((iniopt == 2 .or. iniopt == 12) .and. stime <= 0.0) then
    call fatal('iniopt=2 needs the restart time stime > 0')
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
    dtms = 10_8*int(1.e2*(dtbig+.001e0), 8)
    write(*,'(a,i8)') 'nsteps', 1000_8*int(etime+.1e0, 8)/dtms -        &
     &  1000_8*int(stime+.1e0, 8)/dtms
    write(*,'(a,l2,4i4,2a)') 'outfomat', outf, dmpfmt, dmplev, dmpmon,  &
     &  resopt, ' ', trim(dmpvar(1:15))
    write(*,'(a,3es24.16)') 'intervals', dmpitv, resitv, stime
    write(*,'(a,i4,4a)') 'files', iniopt, ' ', trim(crsdir), ' ',       &
     &  trim(exprim)
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

  opsides = wbc == 7
! fmois (chkmoist.f90:113-155): qv > 0 anywhere on i,j < n-1, k < nk
  allocate(qvb(0:xdim+1,0:ydim+1,1:zdim))
  rc = cress_download(ctx, fid_qv, 1, qvb)
  call check(rc, 'cress_download')
  moist = maxval(qvb(1:xdim-1,1:ydim-1,1:zdim-1)) > 0.0
  deallocate(qvb)
  if(iniopt == 2 .or. iniopt == 12) then
    call read_restart(ctx, trim(out_name(crsdir, exprim, 'res',         &
     &  1000_8*int(stime+.1e0, 8))), xdim, ydim, zdim, cphopt, opsides, dx, &
     &  dy, gwave, dtbig, dtsml, moist)
  end if

! steps ib = nb0..nb1 from ctime = dtms*(ib-1) to ftime = dtms*ib
  dtms = 10_8*int(1.e2*(dtbig+.001e0), 8)
  nb0 = 1000_8*int(stime+.1e0, 8)/dtms + 1
  nb1 = 1000_8*int(etime+.1e0, 8)/dtms
  etms = 1000_8*int(etime+.1e0, 8)
  nsteps = int(max(nb1 - nb0 + 1, 0_8))
  if(outf .and. dtms*(nb0 - 1) == 0) then
    call dump_at(0_8, 0)
  end if
  lastfirst = .false.
  ib = nb0
  do while(ib <= nb1)
    ctime = dtms*(ib - 1)
    first = 0
    if(ctime == 0) first = 1
    if(outf) then
      n = 1
    else
      n = int(nb1 - ib + 1)
      if(first == 1) n = 1
    end if
    rc = cress_large_step(ctx, n, first)
    call check(rc, 'cress_large_step')
    lastfirst = first == 1
    ib = ib + n
    ftime = dtms*(ib - 1)
    if(outf) then
      if(mod(ftime, 1000_8*int(dmpitv+.1e0, 8)) == 0 .or. ftime == etms) then
        call dump_at(ftime, 1)
      end if
      if(resopt == 1) then
        if(mod(ftime, 1000_8*int(resitv+.1e0, 8)) == 0 .or. ftime == etms) then
          rc = cress_sync(ctx)
          call check(rc, 'cress_sync')
          if(lastfirst) then
            call write_restart(ctx, trim(out_name(crsdir, exprim, 'res', ftime)), &
     &        xdim, ydim, zdim, moist, cphopt, opsides, dx, dy, gwave,       &
     &        .5e0*dtbig, .5e0*dtsml)
          else
            call write_restart(ctx, trim(out_name(crsdir, exprim, 'res', ftime)), &
     &        xdim, ydim, zdim, moist, cphopt, opsides, dx, dy, gwave,       &
     &        dtbig, dtsml)
          end if
        end if
      end if
    end if
  end do
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

! The dump of time t (ms) from level lev, then the reset of the dumped
! precipitation accumulations.
  subroutine dump_at(t, lev)
    integer(8), intent(in) :: t
    integer, intent(in) :: lev
    integer(c_int) :: rc2
    rc2 = cress_sync(ctx)
    call check(rc2, 'cress_sync')
    call write_dump(ctx, trim(out_name(crsdir, exprim, 'dmp', t)), xdim, &
     &  ydim, zdim, moist, cphopt, dmpvar, lev)
    if(moist .and. dmpvar(14:14) == 'o') then
      call reset_accum(ctx, xdim, ydim, cphopt)
    end if
  end subroutine dump_at

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
    case('runame')
      read(10, nml=runame, iostat=ios2)
    case('drname')
      read(10, nml=drname, iostat=ios2)
    case('initype')
      read(10, nml=initype, iostat=ios2)
    case('outfomat')
      read(10, nml=outfomat, iostat=ios2)
      outf = ios2 == 0
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
