!----------------------------------------------------------------------!
! aphid_parallel_driver.f90 -- program aphid_gpu_parallel: the          !
! reference's multi-frequency / multi-source job (aphid_parallel,       !
! src/aphid_parallel.f90:40-193; formats in man/aphid.1:168-207) run as !
! one batched solve on one GPU instead of one MPI process per system.   !
!                                                                      !
!   aphid_gpu_parallel [-w dir] [-s]                                   !
!                                                                      !
! Reads dir/aphid_mpi.freq, dir/aphid_mpi.input, dir/aphid_mpi.source  !
! (CS jobs) and dir/aphid.mesh.  System p (0-based) is the reference's  !
! process p: frequency p mod NFREQ; for MT the x- then y-polarised      !
! plane wave (serial modes 2, 3), for CS dipole array p div NFREQ       !
! (mode 5).  Each system's outputs go to dir/tmp-<p+1> (zero-padded to  !
! the digits of the system count, as the reference names its process   !
! directories): the generated aphid.input, qmr.iter (iteration,         !
! residual norm, seconds, as each of the reference's processes writes   !
! it, src/aphid_qmr.f90:381; at most 32M / NSYS lines per system) and   !
! aphid_solution.bin (raw A_x, A_y, A_z, Phi).  With -s the batch       !
! runs as a frequency sweep of three levels: per source, in ascending  !
! order (place q = 0, 1, ...), every fourth frequency starts from      !
! x0 = 0 and each other one from a fit of the solutions d places below !
! and above it, d the lowest set bit of q                              !
! (aphid_batch_set_sweep_guess).                                       !
!----------------------------------------------------------------------!
      program aphid_gpu_parallel
      use iso_c_binding
      use aphid_gpu
      implicit none

      integer, parameter :: dp = 8
      real(dp), parameter :: pi = 3.141592653589793d0

      character(len=256) :: wdir, arg, line, head(3), tail(64), ltol, lmax
      character(len=2)   :: stype
      character(len=1)   :: model
      character(len=8)   :: sdir
      character(len=16)  :: tname
      integer :: nargs, ia, ios, nfreq, nsrc, nsys, p, s, f, q, ntail, nd
      integer :: nx, ny, nz, nblk, b, mode, digits, rc, maxnd, max_it
      integer :: ip, jp, kp
      real(dp) :: tol, secs
      real(dp), allocatable :: freqs(:), xn(:), yn(:), zn(:)
      real(c_float), allocatable, target :: blk(:,:,:,:)
      integer(c_int), allocatable :: ndip(:), dd(:,:), di(:,:), dj(:,:), dk(:,:)
      integer(c_int), allocatable :: skipped(:), ops(:)
      complex(c_double_complex), allocatable :: bx(:,:,:), by(:,:,:)
      complex(c_double_complex), allocatable :: bz(:,:,:), bp(:,:,:)
      type(aphid_sys_report), allocatable :: reps(:)
      type(c_ptr) :: ctx, batch, pa, pb, pc, pd
      logical :: sweep
      integer, allocatable :: order(:)
      integer :: g, nfrom, qz, sd
      integer(c_int) :: from(2)
      real(c_double), allocatable :: hist(:), htime(:)
      integer(c_int) :: hcap, hlen

      wdir = '.'
      sweep = .false.
      nargs = command_argument_count()
      ia = 1
      do while (ia <= nargs)
        call get_command_argument(ia, arg)
        if (arg(1:2) == '-w' .and. ia < nargs) then
          call get_command_argument(ia + 1, wdir)
          ia = ia + 2
        else
          if (arg(1:2) == '-s') sweep = .true.
          ia = ia + 1
        end if
      end do

!     aphid_mpi.freq: 3 comment lines, NFREQ, then the frequencies
      open(unit=11, file=trim(wdir)//'/aphid_mpi.freq', status='old', iostat=ios)
      if (ios /= 0) call die('cannot open aphid_mpi.freq')
      read(11, *); read(11, *); read(11, *)
      read(11, *, iostat=ios) nfreq
      if (ios /= 0 .or. nfreq < 1) call die('reading NFREQ from aphid_mpi.freq')
      allocate(freqs(nfreq))
      do f = 1, nfreq
        read(11, *, iostat=ios) freqs(f)
        if (ios /= 0) call die('reading a frequency from aphid_mpi.freq')
      end do
      close(11)

!     aphid_mpi.input: 3 comment lines, MT/CS, target, max_it, model, ...
      open(unit=12, file=trim(wdir)//'/aphid_mpi.input', status='old', iostat=ios)
      if (ios /= 0) call die('cannot open aphid_mpi.input')
      do q = 1, 3
        read(12, '(a)') head(q)
      end do
      read(12, *) stype
      read(12, '(a)') ltol
      read(ltol, *, iostat=ios) tol
      if (ios /= 0) call die('reading the QMR target from aphid_mpi.input')
      read(12, '(a)') lmax
      read(lmax, *, iostat=ios) max_it
      if (ios /= 0) call die('reading max QMR iterates from aphid_mpi.input')
      ntail = 0
      do
        read(12, '(a)', iostat=ios) line
        if (ios /= 0 .or. ntail == size(tail)) exit
        ntail = ntail + 1
        tail(ntail) = line
      end do
      close(12)
      if (ntail < 1) call die('aphid_mpi.input has no conductivity model line')
      model = adjustl(tail(1))
      if (stype /= 'MT' .and. stype /= 'CS') call die('job type must be MT or CS')

!     aphid_mpi.source (CS): 3 comment lines, NSRC, 1 comment line; per
!     array 3 comment lines, N, N dipoles, 1 separator line
      if (stype == 'MT') then
        nsrc = 2
        maxnd = 1
        allocate(ndip(nsrc), dd(maxnd,nsrc), di(maxnd,nsrc), dj(maxnd,nsrc), dk(maxnd,nsrc))
        ndip = 0
      else
        open(unit=13, file=trim(wdir)//'/aphid_mpi.source', status='old', iostat=ios)
        if (ios /= 0) call die('cannot open aphid_mpi.source')
        read(13, *); read(13, *); read(13, *)
        read(13, *) nsrc
        read(13, *)
        maxnd = 0                              ! first pass: array sizes
        do s = 1, nsrc
          read(13, *); read(13, *); read(13, *)
          read(13, *) nd
          maxnd = max(maxnd, nd)
          do q = 1, nd + 1
            read(13, *)
          end do
        end do
        rewind(13)
        allocate(ndip(nsrc), dd(max(maxnd,1),nsrc), di(max(maxnd,1),nsrc),  &
                 dj(max(maxnd,1),nsrc), dk(max(maxnd,1),nsrc))
        read(13, *); read(13, *); read(13, *); read(13, *); read(13, *)
        do s = 1, nsrc
          read(13, *); read(13, *); read(13, *)
          read(13, *) ndip(s)
          do q = 1, ndip(s)
            read(13, *) sdir, ip, jp, kp
            select case (trim(sdir))
              case ('+X'); dd(q,s) = 1
              case ('-X'); dd(q,s) = -1
              case ('+Y'); dd(q,s) = 2
              case ('-Y'); dd(q,s) = -2
              case ('+Z'); dd(q,s) = 3
              case ('-Z'); dd(q,s) = -3
              case default; dd(q,s) = 0
            end select
            di(q,s) = ip; dj(q,s) = jp; dk(q,s) = kp
          end do
          read(13, *)
        end do
        close(13)
      end if

!     aphid.mesh as for the serial program
      open(unit=14, file=trim(wdir)//'/aphid.mesh', status='old', iostat=ios)
      if (ios /= 0) call die('cannot open aphid.mesh')
      read(14, *) nx, ny, nz
      allocate(xn(nx), yn(ny), zn(nz))
      read(14, *); read(14, *) xn
      read(14, *); read(14, *) yn
      read(14, *); read(14, *) zn
      select case (model)
        case ('R', 'I'); nblk = 1
        case ('C'); nblk = 2
        case ('D'); nblk = 4
        case default; call die('conductivity model must be R, C, D or I')
      end select
      allocate(blk(nx-1, ny-1, nz-1, nblk))
      do b = 1, nblk
        read(14, *)
        read(14, *) blk(:, :, :, b)
      end do
      close(14)
      pa = c_null_ptr; pb = c_null_ptr; pc = c_null_ptr; pd = c_null_ptr
      select case (model)
        case ('R'); pa = c_loc(blk(1,1,1,1))
        case ('C'); pa = c_loc(blk(1,1,1,1)); pb = c_loc(blk(1,1,1,2))
        case ('D'); pa = c_loc(blk(1,1,1,1)); pb = c_loc(blk(1,1,1,2))
                    pc = c_loc(blk(1,1,1,3)); pd = c_loc(blk(1,1,1,4))
        case ('I'); pb = c_loc(blk(1,1,1,1))
      end select

!     one operator per frequency, then every system in process order
      nsys = nsrc * nfreq
      if (nsys > 999) call die('more than 999 systems (the reference names tmp-NNN at most)')
      rc = aphid_batch_create(batch, -1_c_int)
      call check(rc, 'aphid_batch_create')
      allocate(ops(nfreq))
      do f = 1, nfreq
        ops(f) = aphid_batch_add_operator(batch, int(nx, c_int), int(ny, c_int),      &
                   int(nz, c_int), xn, yn, zn, model, 2.0_dp * pi * freqs(f), pa, pb, pc, pd)
        call check(ops(f), 'aphid_batch_add_operator')
      end do
      rc = aphid_ctx_create(ctx, -1_c_int)
      call check(rc, 'aphid_ctx_create')
      rc = aphid_set_mesh(ctx, int(nx, c_int), int(ny, c_int), int(nz, c_int), xn, yn, zn)
      call check(rc, 'aphid_set_mesh')
      allocate(bx(nx-1,ny,nz), by(nx,ny-1,nz), bz(nx,ny,nz-1), bp(nx,ny,nz))
      allocate(skipped(max(maxnd,1)))
      do p = 0, nsys - 1
        s = p / nfreq + 1
        f = mod(p, nfreq) + 1
        if (stype == 'MT') then
          mode = 1 + s
          rc = aphid_build_rhs(ctx, int(mode, c_int), 0_c_int, dd(:,1), di(:,1),   &
                               dj(:,1), dk(:,1), skipped, bx, by, bz, bp)
        else
          mode = 5
          rc = aphid_build_rhs(ctx, 5_c_int, int(ndip(s), c_int), dd(:,s), di(:,s),  &
                               dj(:,s), dk(:,s), skipped, bx, by, bz, bp)
        end if
        call check(rc, 'aphid_build_rhs')
        rc = aphid_batch_add_system(batch, ops(f), bx, by, bz, bp, c_null_ptr,       &
                                    c_null_ptr, c_null_ptr, c_null_ptr)
        call check(rc, 'aphid_batch_add_system')
      end do
      rc = aphid_ctx_destroy(ctx)

!     frequency sweep: per source, the frequencies in ascending order
!     (place qz = f - 1); every fourth from x0 = 0, each other one from
!     the solutions sd places below and above it (same source), sd the
!     lowest set bit of qz
      if (sweep .and. nfreq >= 2) then
        allocate(order(nfreq))
        do f = 1, nfreq
          order(f) = f
        end do
        do f = 2, nfreq
          q = order(f)
          g = f - 1
          do while (g >= 1)
            if (freqs(order(g)) <= freqs(q)) exit
            order(g + 1) = order(g)
            g = g - 1
          end do
          order(g + 1) = q
        end do
        do s = 1, nsys / nfreq
          do f = 2, nfreq
            qz = f - 1
            if (mod(qz, 4) == 0) cycle
            sd = iand(qz, -qz)
            nfrom = 1
            from(1) = int((s - 1) * nfreq + order(f - sd) - 1, c_int)
            if (f + sd <= nfreq) then
              nfrom = 2
              from(2) = int((s - 1) * nfreq + order(f + sd) - 1, c_int)
            end if
            rc = aphid_batch_set_sweep_guess(batch,                       &
                   int((s - 1) * nfreq + order(f) - 1, c_int),            &
                   int(nfrom, c_int), from)
            call check(rc, 'aphid_batch_set_sweep_guess')
          end do
        end do
      end if

!     residual histories for qmr.iter: one per system, 256 MB at most
      hcap = int(max(1, min(max_it, 33554432 / nsys)), c_int)
      allocate(hist(hcap), htime(hcap))
      rc = aphid_batch_set_history(batch, hcap)
      call check(rc, 'aphid_batch_set_history')

      allocate(reps(nsys))
      rc = aphid_batch_solve(batch, tol, int(max_it, c_int), 0_c_int, reps, secs)
      call check(rc, 'aphid_batch_solve')
      write(6, '(a,i6,a,f10.3,a)') ' Solved ', nsys, ' systems in one batch: ', secs, ' s'

      digits = 1
      if (nsys > 9) digits = 2
      if (nsys > 99) digits = 3
      do p = 0, nsys - 1
        s = p / nfreq + 1
        f = mod(p, nfreq) + 1
        mode = 5
        if (stype == 'MT') mode = 1 + s
        write(tname, '(a,i0.'//char(48 + digits)//')') 'tmp-', p + 1
        call execute_command_line('mkdir -p "'//trim(wdir)//'/'//trim(tname)//'"')
!       the per-process input the reference generates (:146-162): the
!       target and max-iteration lines are copied as they stand
        open(unit=21, file=trim(wdir)//'/'//trim(tname)//'/aphid.input', status='replace')
        write(21, '(a)') trim(head(1))
        write(21, '(a)') 'APhiD input file generated from ../aphid_mpi.input'
        write(21, '(a)') trim(head(3))
        write(21, '(i4,2x,a)') mode, ' !! operational mode'
        write(21, '(a)') trim(ltol)
        write(21, '(a)') trim(lmax)
        write(21, '(e14.7,2x,a)') freqs(f), ' !! frequency [Hz]'
        do q = 1, ntail
          write(21, '(a)') trim(tail(q))
        end do
        close(21)
        if (mode == 5) then
          open(unit=22, file=trim(wdir)//'/'//trim(tname)//'/aphid.source', status='replace')
          do q = 1, ndip(s)
            write(22, '(a,3(1x,i0))') dirname(dd(q,s)), di(q,s), dj(q,s), dk(q,s)
          end do
          close(22)
        end if
        rc = aphid_batch_get_history(batch, int(p, c_int), hist, htime, hcap, hlen)
        call check(rc, 'aphid_batch_get_history')
        open(unit=23, file=trim(wdir)//'/'//trim(tname)//'/qmr.iter', status='replace')
        do q = 1, hlen
          write(23, '(i8,1x,es24.16,1x,es24.16)') q, hist(q), htime(q)
        end do
        close(23)
        rc = aphid_batch_get_solution(batch, int(p, c_int), bx, by, bz, bp)
        call check(rc, 'aphid_batch_get_solution')
        open(unit=24, file=trim(wdir)//'/'//trim(tname)//'/aphid_solution.bin',       &
             access='stream', form='unformatted', status='replace')
        write(24) bx, by, bz, bp
        close(24)
        write(6, '(1x,a,f14.6,a,i3,a,i7,a,es10.3)') trim(tname)//': f = ', freqs(f),   &
              ' Hz, mode ', mode, ', iterations ', reps(p+1)%iters,                   &
              ', true rel. residual ', reps(p+1)%true_rel_res
      end do
      rc = aphid_batch_destroy(batch)

      contains

      function dirname(code) result(nm)
        integer(c_int), intent(in) :: code
        character(len=2) :: nm
        select case (code)
          case (1); nm = '+X'
          case (-1); nm = '-X'
          case (2); nm = '+Y'
          case (-2); nm = '-Y'
          case (3); nm = '+Z'
          case default; nm = '-Z'
        end select
      end function dirname

      subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(*), intent(in) :: what
        if (code < 0) then
          write(6, '(a)') ' >> ERROR in '//what//': '//trim(aphid_error_string())
          stop 1
        end if
      end subroutine check

      subroutine die(msg)
        character(*), intent(in) :: msg
        write(6, '(a)') ' >> ERROR '//msg
        stop 1
      end subroutine die

      end program aphid_gpu_parallel
