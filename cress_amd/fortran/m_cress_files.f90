!-----------------------------------------------------------------------
! m_cress_files -- the reference's restart and dump files for the state
! held by a device context (m_cress_hip), one sub-domain and one group.
!
! Restart (outres.f90:886-1163, read back by rdres.f90 in the same order):
! sequential unformatted, one record per statement, arrays with their full
! halo, for sfcopt=0, advopt<=3, aslopt=trkopt=tubopt=diaopt=0 and
! cphopt 0 or 2:
!   fmois(1:5);  ubr, vbr, pbr, ptbr, qvbr;  u, v, w, pp, ptp (present);
!   up, vp, wp, ppp, ptpp (past);
!   open sides: ucpx, ucpy, vcpx, vcpy, wcpx, wcpy, pcpx, pcpy, ptcpx, ptcpy;
!   moist: qv, qvp;  moist + open: qvcpx, qvcpy;
!   moist, cphopt>=1: qwtr(qc,qr), qwtrp;  + open: qwcpx, qwcpy;
!   moist, cphopt=2: qice(qi,qs,qg), nice(nci), qicep, nicep;
!     + open: qicpx, qicpy, nicpx, nicpy;
!   moist, cphopt>=1: prwtr(:,:,1,1:2), prwtr(:,:,2,1:2);
!   moist, cphopt=2: price(:,:,1,1:3), price(:,:,2,1:3).
!
! Dump, dmplev=1, dmpfmt=2, dmpmon=0 (outdmp.f90:385-700, 1506-1590,
! outdmp3d.f90, outdmp2d.f90): direct access, one record per level
! k=2..nk-2 of ((var(i,j,k), i=2,ni-2), j=2,nj-2), u, v and w averaged to
! the scalar points, in the order and with the flags of dmpvar; then the
! surface precipitation (dmpvar(14:14)), whose accumulations are reset to
! zero once written.
!
! File names (outres.f90:964-972, opendmp.f90:566-573 for one group):
! <crsdir><exprim>res<ctime in s, i8.8>.pe0000.bin, and dmp for dumps.
!-----------------------------------------------------------------------
module m_cress_files

  use iso_c_binding
  use m_cress_hip

  implicit none

  private
  public :: out_name, write_restart, read_restart, write_dump,         &
     &      phase_pair, reset_accum

contains

  function out_name(crsdir, exprim, kind, ctime) result(fn)
    character(len=*), intent(in) :: crsdir, exprim, kind
    integer(8), intent(in) :: ctime        ! ms
    character(len=512) :: fn
    character(len=22) :: tail
    write(tail, '(a3,i8.8,a3,i4.4,a4)') kind, ctime/1000_8, '.pe', 0, '.bin'
    fn = trim(crsdir)//trim(exprim)//tail
  end function out_name

! Phase speeds of the open sides as s_phasev leaves them after a step with
! big and small time steps dtb, dts (option 7, mfcopt=0; arrays zero where
! it does not write them, allocslv.f90:2015).  kind 1 u (phvuvw.f90:437-
! 808: west set and clipped on k=2..nk-1 but scaled on k=2..nk-2), 2 v,
! 3 w, 4 p (phvs 'sml': scaled by dts/dtb), 5 the large-step scalars
! (phvs 'big', phvs.f90:375-460: unscaled).
  subroutine phase_pair(kind, ni, nj, nk, dx, dy, gwave, dtb, dts, cx, cy)
    integer, intent(in) :: kind, ni, nj, nk
    real, intent(in) :: dx, dy, gwave, dtb, dts
    real, intent(out) :: cx(1:nj,1:nk,1:2), cy(1:ni,1:nk,1:2)
    real :: dxiv, dyiv, gx, gy, gxn, gyn, sc, w, e, s, n
    integer :: j, k, khi, jhi, ihi
    dxiv = 1.e0/dx
    dyiv = 1.e0/dy
    gx = gwave*dtb*dxiv
    gy = gwave*dtb*dyiv
    gxn = -(gwave*dtb*dxiv)
    gyn = -(gwave*dtb*dyiv)
    sc = dts/dtb
    cx = 0.e0
    cy = 0.e0
    w = max(gxn, -1.e0)
    e = min(gx, 1.e0)
    s = max(gyn, -1.e0)
    n = min(gy, 1.e0)
    if(kind /= 5) then
      w = w*sc
      e = e*sc
      s = s*sc
      n = n*sc
    end if
    khi = nk - 2
    if(kind == 3) khi = nk - 1
    jhi = nj - 1
    if(kind == 2) jhi = nj
    ihi = ni - 1
    if(kind == 1) ihi = ni
    do k = 2, khi
      do j = 1, jhi
        cx(j,k,1) = w
        cx(j,k,2) = e
      end do
      cy(1:ihi,k,1) = s
      cy(1:ihi,k,2) = n
    end do
    if(kind == 1) then
      do j = 1, nj - 1
        cx(j,nk-1,1) = max(gxn, -1.e0)
      end do
    end if
  end subroutine phase_pair

  subroutine get3(ctx, fid, lev, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: fid
    integer, intent(in) :: lev
    real(c_float), intent(inout) :: a(*)
    call chk(cress_download(ctx, fid, int(lev, c_int), a), 'cress_download')
  end subroutine get3

  subroutine set3(ctx, fid, lev, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: fid
    integer, intent(in) :: lev
    real(c_float), intent(in) :: a(*)
    call chk(cress_upload(ctx, fid, int(lev, c_int), a), 'cress_upload')
  end subroutine set3

  subroutine write_restart(ctx, fn, ni, nj, nk, moist, cphopt, opn,    &
     &                     dx, dy, gwave, dtb, dts)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: fn
    integer, intent(in) :: ni, nj, nk, cphopt
    logical, intent(in) :: moist, opn
    real, intent(in) :: dx, dy, gwave, dtb, dts
    real(c_float), allocatable :: a(:,:,:), b(:,:,:), c3(:,:,:)
    real(c_float), allocatable :: pw(:,:,:,:), pi(:,:,:,:)
    real, allocatable :: cx(:,:,:), cy(:,:,:)
    integer :: u, ios, n, kind
    integer(c_int) :: f
    character(len=5) :: fmois

    allocate(a(0:ni+1,0:nj+1,1:nk), b(0:ni+1,0:nj+1,1:nk),              &
     &       c3(0:ni+1,0:nj+1,1:nk))
    allocate(cx(1:nj,1:nk,1:2), cy(1:ni,1:nk,1:2))
    open(newunit=u, file=trim(fn), access='sequential', form='unformatted', &
     &   status='replace', action='write', iostat=ios)
    if(ios /= 0) call die('cannot create '//trim(fn))
    fmois = 'dry  '
    if(moist) fmois = 'moist'
    write(u) fmois
    do f = fid_ubr, fid_qvbr
      call get3(ctx, f, 0, a)
      write(u) a
    end do
    do f = fid_u, fid_ptp
      call get3(ctx, f, 1, a)
      write(u) a
    end do
    do f = fid_u, fid_ptp
      call get3(ctx, f, 0, a)
      write(u) a
    end do
    if(opn) then
      do kind = 1, 5
        call phase_pair(kind, ni, nj, nk, dx, dy, gwave, dtb, dts, cx, cy)
        write(u) cx
        write(u) cy
      end do
    end if
    call phase_pair(5, ni, nj, nk, dx, dy, gwave, dtb, dts, cx, cy)
    if(moist) then
      call get3(ctx, fid_qv, 1, a)
      write(u) a
      call get3(ctx, fid_qv, 0, a)
      write(u) a
      if(opn) then
        write(u) cx
        write(u) cy
      end if
    end if
    if(moist .and. cphopt >= 1) then
      do n = 1, 0, -1
        call get3(ctx, fid_qc, n, a)
        call get3(ctx, fid_qr, n, b)
        write(u) a, b
      end do
      if(opn) then
        write(u) cx, cx
        write(u) cy, cy
      end if
    end if
    if(moist .and. cphopt == 2) then
      do n = 1, 0, -1
        call get3(ctx, fid_qi, n, a)
        call get3(ctx, fid_qs, n, b)
        call get3(ctx, fid_qg, n, c3)
        write(u) a, b, c3
        call get3(ctx, fid_nci, n, a)
        write(u) a
      end do
      if(opn) then
        write(u) cx, cx, cx
        write(u) cy, cy, cy
        write(u) cx
        write(u) cy
      end if
    end if
    if(moist .and. cphopt >= 1) then
      allocate(pw(0:ni+1,0:nj+1,1:2,1:2))
      call get3(ctx, fid_prwtr, 0, pw)
      write(u) (pw(:,:,1,n), n = 1, 2)
      write(u) (pw(:,:,2,n), n = 1, 2)
    end if
    if(moist .and. cphopt == 2) then
      allocate(pi(0:ni+1,0:nj+1,1:2,1:3))
      call get3(ctx, fid_price, 0, pi)
      write(u) (pi(:,:,1,n), n = 1, 3)
      write(u) (pi(:,:,2,n), n = 1, 3)
    end if
    close(u)
  end subroutine write_restart

! Reads a restart file written for the same options and uploads its
! levels; the phase speeds it holds must be those of a full step (a file
! written after the cold-start step holds the speeds of half steps, which
! the next call cannot take).
  subroutine read_restart(ctx, fn, ni, nj, nk, cphopt, opn, dx, dy,    &
     &                    gwave, dtbig, dtsml, moist)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: fn
    integer, intent(in) :: ni, nj, nk, cphopt
    logical, intent(in) :: opn
    real, intent(in) :: dx, dy, gwave, dtbig, dtsml
    logical, intent(out) :: moist
    real(c_float), allocatable :: a(:,:,:), b(:,:,:), c3(:,:,:)
    real(c_float), allocatable :: pw(:,:,:,:), pi(:,:,:,:)
    real, allocatable :: cx(:,:,:), cy(:,:,:), rx(:,:,:), ry(:,:,:)
    real, allocatable :: rx2(:,:,:,:), ry2(:,:,:,:)
    integer :: u, ios, n, kind, m
    integer(c_int) :: f
    character(len=5) :: fmois

    allocate(a(0:ni+1,0:nj+1,1:nk), b(0:ni+1,0:nj+1,1:nk),              &
     &       c3(0:ni+1,0:nj+1,1:nk))
    allocate(cx(1:nj,1:nk,1:2), cy(1:ni,1:nk,1:2))
    allocate(rx(1:nj,1:nk,1:2), ry(1:ni,1:nk,1:2))
    open(newunit=u, file=trim(fn), access='sequential', form='unformatted', &
     &   status='old', action='read', iostat=ios)
    if(ios /= 0) call die('cannot open restart file '//trim(fn))
    read(u, iostat=ios) fmois
    if(ios /= 0) call die('not a restart file: '//trim(fn))
    moist = fmois == 'moist'
    if(.not. moist .and. fmois /= 'dry  ') call die('not a restart file: '//trim(fn))
    do f = fid_ubr, fid_qvbr
      call rd(a)
      call set3(ctx, f, 0, a)
    end do
    do f = fid_u, fid_ptp
      call rd(a)
      call set3(ctx, f, 1, a)
    end do
    do f = fid_u, fid_ptp
      call rd(a)
      call set3(ctx, f, 0, a)
    end do
    if(opn) then
      do kind = 1, 5
        call phase_pair(kind, ni, nj, nk, dx, dy, gwave, dtbig, dtsml, cx, cy)
        read(u, iostat=ios) rx
        if(ios == 0) read(u, iostat=ios) ry
        if(ios /= 0) call die('truncated restart file')
        call same_speeds(cx, cy, rx, ry)
      end do
    end if
    call phase_pair(5, ni, nj, nk, dx, dy, gwave, dtbig, dtsml, cx, cy)
    if(moist) then
      call rd(a)
      call set3(ctx, fid_qv, 1, a)
      call rd(a)
      call set3(ctx, fid_qv, 0, a)
      if(opn) call rd_speeds(1)
    end if
    if(moist .and. cphopt >= 1) then
      do n = 1, 0, -1
        read(u, iostat=ios) a, b
        if(ios /= 0) call die('truncated restart file')
        call set3(ctx, fid_qc, n, a)
        call set3(ctx, fid_qr, n, b)
      end do
      if(opn) call rd_speeds(2)
    end if
    if(moist .and. cphopt == 2) then
      do n = 1, 0, -1
        read(u, iostat=ios) a, b, c3
        if(ios /= 0) call die('truncated restart file')
        call set3(ctx, fid_qi, n, a)
        call set3(ctx, fid_qs, n, b)
        call set3(ctx, fid_qg, n, c3)
        call rd(a)
        call set3(ctx, fid_nci, n, a)
      end do
      if(opn) then
        call rd_speeds(3)
        call rd_speeds(1)
      end if
    end if
    if(moist .and. cphopt >= 1) then
      allocate(pw(0:ni+1,0:nj+1,1:2,1:2))
      read(u, iostat=ios) (pw(:,:,1,n), n = 1, 2)
      if(ios == 0) read(u, iostat=ios) (pw(:,:,2,n), n = 1, 2)
      if(ios /= 0) call die('truncated restart file')
      call set3(ctx, fid_prwtr, 0, pw)
    end if
    if(moist .and. cphopt == 2) then
      allocate(pi(0:ni+1,0:nj+1,1:2,1:3))
      read(u, iostat=ios) (pi(:,:,1,n), n = 1, 3)
      if(ios == 0) read(u, iostat=ios) (pi(:,:,2,n), n = 1, 3)
      if(ios /= 0) call die('truncated restart file')
      call set3(ctx, fid_price, 0, pi)
    end if
    close(u)

  contains

    subroutine rd(x)
      real(c_float), intent(out) :: x(0:ni+1,0:nj+1,1:nk)
      integer :: ios2
      read(u, iostat=ios2) x
      if(ios2 /= 0) call die('truncated restart file')
    end subroutine rd

    ! m species' copies of the large-step scalar pair
    subroutine rd_speeds(nm)
      integer, intent(in) :: nm
      integer :: ios2
      allocate(rx2(1:nj,1:nk,1:2,1:nm), ry2(1:ni,1:nk,1:2,1:nm))
      read(u, iostat=ios2) rx2
      if(ios2 == 0) read(u, iostat=ios2) ry2
      if(ios2 /= 0) call die('truncated restart file')
      do m = 1, nm
        call same_speeds(cx, cy, rx2(:,:,:,m), ry2(:,:,:,m))
      end do
      deallocate(rx2, ry2)
    end subroutine rd_speeds

    subroutine same_speeds(px, py, qx, qy)
      real, intent(in) :: px(:,:,:), py(:,:,:), qx(:,:,:), qy(:,:,:)
      if(any(px /= qx) .or. any(py /= qy)) then
        call die('restart phase speeds are not those of a full step '//&
     &           '(a file written after the cold-start step)')
      end if
    end subroutine same_speeds

  end subroutine read_restart

! Writes the dump of one time: level lev of the prognostics (the past
! level 0 at the forecast start, as slvdrv.f90:1243-1249 passes up, vp,
! ...; the present level 1 after a step).
  subroutine write_dump(ctx, fn, ni, nj, nk, moist, cphopt, dmpvar, lev)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: fn, dmpvar
    integer, intent(in) :: ni, nj, nk, cphopt, lev
    logical, intent(in) :: moist
    real(c_float), allocatable :: a(:,:,:), b(:,:,:)
    real(c_float), allocatable :: pw(:,:,:,:), pi(:,:,:,:)
    integer :: u, ios, rl, rec, i, j, k, n
    integer(c_int) :: qf(5)

    allocate(a(0:ni+1,0:nj+1,1:nk), b(0:ni+1,0:nj+1,1:nk))
    inquire(iolength=rl) a(0,0,1)
    open(newunit=u, file=trim(fn), access='direct', form='unformatted', &
     &   recl=(ni-3)*(nj-3)*rl, status='replace', action='write', iostat=ios)
    if(ios /= 0) call die('cannot create '//trim(fn))
    rec = 0
    if(dmpvar(1:1) == 'o') then
      call get3(ctx, fid_ubr, 0, a)
      call put3('oxx', a)
    end if
    if(dmpvar(1:1) == 'o' .or. dmpvar(1:1) == '-') then
      call get3(ctx, fid_u, lev, a)
      call put3('oxx', a)
    end if
    if(dmpvar(2:2) == 'o') then
      call get3(ctx, fid_vbr, 0, a)
      call put3('xox', a)
    end if
    if(dmpvar(2:2) == 'o' .or. dmpvar(2:2) == '-') then
      call get3(ctx, fid_v, lev, a)
      call put3('xox', a)
    end if
    if(dmpvar(3:3) == 'o') then
      call get3(ctx, fid_w, lev, a)
      call put3('xxo', a)
    end if
    call pair(4, fid_pbr, fid_pp)
    call pair(5, fid_ptbr, fid_ptp)
    if(moist) then
      if(dmpvar(6:6) == 'o') then
        call get3(ctx, fid_qvbr, 0, a)
        call put3('xxx', a)
      end if
      if(dmpvar(6:6) == 'o' .or. dmpvar(6:6) == '-') then
        call get3(ctx, fid_qv, lev, a)
        call put3('xxx', a)
      end if
      if(cphopt >= 1 .and. dmpvar(7:7) == 'o') then
        qf = (/ fid_qc, fid_qr, fid_qi, fid_qs, fid_qg /)
        do n = 1, 5
          if(n >= 3 .and. cphopt < 2) exit
          call get3(ctx, qf(n), lev, a)
          call put3('xxx', a)
        end do
      end if
      if(cphopt >= 1 .and. dmpvar(14:14) == 'o') then
        allocate(pw(0:ni+1,0:nj+1,1:2,1:2))
        call get3(ctx, fid_prwtr, 0, pw)
        do n = 1, 2
          call put2(pw(:,:,1,n))
          call put2(pw(:,:,2,n))
        end do
      end if
      if(cphopt == 2 .and. dmpvar(14:14) == 'o') then
        allocate(pi(0:ni+1,0:nj+1,1:2,1:3))
        call get3(ctx, fid_price, 0, pi)
        do n = 1, 3
          call put2(pi(:,:,1,n))
          call put2(pi(:,:,2,n))
        end do
      end if
    end if
    close(u)

  contains

    ! 'o': base state then perturbation; '-': their sum (totals.f90)
    subroutine pair(pos, fbar, fprt)
      integer, intent(in) :: pos
      integer(c_int), intent(in) :: fbar, fprt
      if(dmpvar(pos:pos) == 'o') then
        call get3(ctx, fbar, 0, a)
        call put3('xxx', a)
        call get3(ctx, fprt, lev, a)
        call put3('xxx', a)
      else if(dmpvar(pos:pos) == '-') then
        call get3(ctx, fbar, 0, b)
        call get3(ctx, fprt, lev, a)
        do k = 1, nk - 1
          do j = 0, nj
            do i = 0, ni
              a(i,j,k) = b(i,j,k) + a(i,j,k)
            end do
          end do
        end do
        call put3('xxx', a)
      end if
    end subroutine pair

    subroutine put3(xo, v)
      character(len=3), intent(in) :: xo
      real(c_float), intent(in) :: v(0:ni+1,0:nj+1,1:nk)
      integer :: ios2
      do k = 2, nk - 2
        rec = rec + 1
        select case(xo)
        case('oxx')
          write(u, rec=rec, iostat=ios2)                                &
     &      ((.5e0*(v(i,j,k)+v(i+1,j,k)), i = 2, ni-2), j = 2, nj-2)
        case('xox')
          write(u, rec=rec, iostat=ios2)                                &
     &      ((.5e0*(v(i,j,k)+v(i,j+1,k)), i = 2, ni-2), j = 2, nj-2)
        case('xxo')
          write(u, rec=rec, iostat=ios2)                                &
     &      ((.5e0*(v(i,j,k)+v(i,j,k+1)), i = 2, ni-2), j = 2, nj-2)
        case default
          write(u, rec=rec, iostat=ios2) ((v(i,j,k), i = 2, ni-2), j = 2, nj-2)
        end select
        if(ios2 /= 0) call die('cannot write '//trim(fn))
      end do
    end subroutine put3

    subroutine put2(v)
      real(c_float), intent(in) :: v(0:ni+1,0:nj+1)
      integer :: ios2
      rec = rec + 1
      write(u, rec=rec, iostat=ios2) ((v(i,j), i = 2, ni-2), j = 2, nj-2)
      if(ios2 /= 0) call die('cannot write '//trim(fn))
    end subroutine put2

  end subroutine write_dump

! The accumulated precipitation is set to zero once dumped
! (outdmp.f90:1538-1580).
  subroutine reset_accum(ctx, ni, nj, cphopt)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: ni, nj, cphopt
    real(c_float), allocatable :: pw(:,:,:,:), pi(:,:,:,:)
    if(cphopt >= 1) then
      allocate(pw(0:ni+1,0:nj+1,1:2,1:2))
      call get3(ctx, fid_prwtr, 0, pw)
      pw(:,:,2,:) = 0.e0
      call set3(ctx, fid_prwtr, 0, pw)
    end if
    if(cphopt == 2) then
      allocate(pi(0:ni+1,0:nj+1,1:2,1:3))
      call get3(ctx, fid_price, 0, pi)
      pi(:,:,2,:) = 0.e0
      call set3(ctx, fid_price, 0, pi)
    end if
  end subroutine reset_accum

  subroutine chk(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if(code /= cress_ok) call die(what//': '//trim(cress_errmsg(code)))
  end subroutine chk

  subroutine die(msg)
    character(len=*), intent(in) :: msg
    write(*,'(2a)') 'cress_run: ', msg
    stop 1
  end subroutine die

end module m_cress_files
