!----------------------------------------------------------------------!
! aphid_driver.f90 -- program aphid_gpu_serial: the reference's serial  !
! flow (aphid_main, src/aphid.f90:6-118) with the solve on the GPU.     !
!                                                                      !
!   aphid_gpu_serial [-w dir] [-r]                                     !
!                                                                      !
! -r: reproducible solve (aphid_set_reproducible): the iterates do not  !
! depend on the stencil launch shape the autotuner picks.               !
! Reads dir/aphid.input, dir/aphid.mesh and (mode 5) dir/aphid.source  !
! in the reference's formats (man/aphid.1:63-124), sets the mesh and    !
! the FP32 cell parameters on the device (read_mesh, :269-454), builds  !
! the right-hand side (gen_rhs, :534-743), and runs the Jacobi-QMR      !
! solve (qmr, src/aphid_qmr.f90:5-434) through the C ABI.  Writes       !
! dir/qmr.iter (iteration, residual norm) as the reference does, the    !
! solution (A_x, A_y, A_z, Phi in the reference's shapes) as a raw      !
! stream file dir/aphid_solution.bin, and the reference's VTK files     !
! (module aphid_vtk): sigma.vtk, aphid_a_phi.vtk, Bx..Ez.vtk and        !
! EM_grid.vtk from the fields computed on the device (gen_fields).      !
!----------------------------------------------------------------------!
      program aphid_gpu_serial
      use iso_c_binding
      use aphid_gpu
      use aphid_vtk
      implicit none

      integer, parameter :: dp = 8
      real(dp), parameter :: pi = 3.141592653589793d0

      character(len=256) :: wdir, arg
      character(len=1)   :: model
      character(len=8)   :: sdir
      integer :: nargs, ia, mode, max_it, nx, ny, nz, ios, ns, s, rc
      integer :: i, ip, jp, kp, nsmax
      real(dp) :: target, freq, omega
      real(dp), allocatable :: xn(:), yn(:), zn(:)
      real(c_float), allocatable, target :: blk(:,:,:,:)
      integer(c_int), allocatable :: sd(:), si(:), sj(:), sk(:), skipped(:)
      complex(c_double_complex), allocatable :: bx(:,:,:), by(:,:,:)
      complex(c_double_complex), allocatable :: bz(:,:,:), bp(:,:,:)
      real(c_double), allocatable, target :: hist(:), htime(:)
      complex(c_double_complex), allocatable :: sig(:,:,:)
      complex(c_double_complex), allocatable :: fbx(:,:,:), fby(:,:,:)
      complex(c_double_complex), allocatable :: fbz(:,:,:), fex(:,:,:)
      complex(c_double_complex), allocatable :: fey(:,:,:), fez(:,:,:)
      complex(c_double_complex), allocatable :: k2x(:,:,:), k2y(:,:,:)
      complex(c_double_complex), allocatable :: k2z(:,:,:), k2p(:,:,:)
      type(c_ptr) :: ctx, pa, pb, pc, pd
      type(aphid_qmr_report) :: rep
      integer :: nblk, b
      logical :: repro

!     -w <dir> (src/aphid.f90:47-60); -r reproducible solve
      wdir = '.'
      repro = .false.
      nargs = command_argument_count()
      ia = 1
      do while (ia <= nargs)
        call get_command_argument(ia, arg)
        if (arg(1:2) == '-w' .and. ia < nargs) then
          call get_command_argument(ia + 1, wdir)
          ia = ia + 2
        else
          if (arg(1:2) == '-r') repro = .true.
          ia = ia + 1
        end if
      end do

!     aphid.input: 3 comment lines, mode, target, max_it, freq, model
      open(unit=11, file=trim(wdir)//'/aphid.input', status='old',      &
           iostat=ios)
      if (ios /= 0) call die('cannot open aphid.input')
      read(11, *)
      read(11, *)
      read(11, *)
      read(11, *, iostat=ios) mode
      if (ios /= 0) call die('reading operational mode from aphid.input')
      read(11, *, iostat=ios) target
      if (ios /= 0) call die('reading qmr_target from aphid.input')
      read(11, *, iostat=ios) max_it
      if (ios /= 0) call die('reading qmr_max_it from aphid.input')
      read(11, *) freq
      read(11, *) model
      close(11)
      if (target < 0.0_dp) call die('invalid target residual: must be positive')
      if (max_it < 0) call die('invalid qmr_max_it: must be positive')
      omega = 2.0_dp * pi * freq

!     aphid.mesh: nodes, three coordinate blocks, model-dependent cell
!     blocks read as real(4)
      open(unit=12, file=trim(wdir)//'/aphid.mesh', status='old',       &
           iostat=ios)
      if (ios /= 0) call die('cannot open aphid.mesh')
      read(12, *) nx, ny, nz
      allocate(xn(nx), yn(ny), zn(nz))
      read(12, *)
      read(12, *) xn
      read(12, *)
      read(12, *) yn
      read(12, *)
      read(12, *) zn
      select case (model)
        case ('R'); nblk = 1
        case ('C'); nblk = 2
        case ('D'); nblk = 4
        case ('I'); nblk = 1
        case default; call die('conductivity model must be R, C, D or I')
      end select
      allocate(blk(nx-1, ny-1, nz-1, nblk))
      do b = 1, nblk
        read(12, *)
        read(12, *) blk(:, :, :, b)
      end do
      close(12)

      rc = aphid_ctx_create(ctx, -1_c_int)
      call check(rc, 'aphid_ctx_create')
      if (repro) then
        rc = aphid_set_reproducible(ctx, 1_c_int)
        call check(rc, 'aphid_set_reproducible')
      end if
      rc = aphid_set_mesh(ctx, int(nx, c_int), int(ny, c_int),          &
                          int(nz, c_int), xn, yn, zn)
      call check(rc, 'aphid_set_mesh')
      pa = c_null_ptr; pb = c_null_ptr; pc = c_null_ptr; pd = c_null_ptr
      select case (model)
        case ('R'); pa = c_loc(blk(1,1,1,1))
        case ('C'); pa = c_loc(blk(1,1,1,1)); pb = c_loc(blk(1,1,1,2))
        case ('D'); pa = c_loc(blk(1,1,1,1)); pb = c_loc(blk(1,1,1,2))
                    pc = c_loc(blk(1,1,1,3)); pd = c_loc(blk(1,1,1,4))
        case ('I'); pb = c_loc(blk(1,1,1,1))
      end select
      rc = aphid_set_cells(ctx, model, omega, pa, pb, pc, pd)
      call check(rc, 'aphid_set_cells')

!     sigma.vtk right after the model is read (src/aphid.f90:79)
      allocate(sig(nx-1, ny-1, nz-1))
      rc = aphid_get_sigma(ctx, sig)
      call check(rc, 'aphid_get_sigma')
      call vtk_sigma(trim(wdir)//'/sigma.vtk', xn, yn, zn, sig)

!     sources (mode 5: one "+X i j k" per line until end of file)
      ns = 0
      nsmax = 0
      if (mode == 5) then
        open(unit=13, file=trim(wdir)//'/aphid.source', status='old',   &
             iostat=ios)
        if (ios /= 0) call die('cannot open aphid.source')
        do
          read(13, *, iostat=ios) sdir, ip, jp, kp
          if (ios /= 0) exit
          nsmax = nsmax + 1
        end do
        rewind(13)
      end if
      allocate(sd(max(nsmax,1)), si(max(nsmax,1)), sj(max(nsmax,1)),   &
               sk(max(nsmax,1)), skipped(max(nsmax,1)))
      if (mode == 5) then
        do s = 1, nsmax
          read(13, *) sdir, ip, jp, kp
          ns = ns + 1
          select case (trim(sdir))
            case ('+X'); sd(ns) = 1
            case ('-X'); sd(ns) = -1
            case ('+Y'); sd(ns) = 2
            case ('-Y'); sd(ns) = -2
            case ('+Z'); sd(ns) = 3
            case ('-Z'); sd(ns) = -3
            case default; sd(ns) = 0
          end select
          si(ns) = ip; sj(ns) = jp; sk(ns) = kp
        end do
        close(13)
      end if

      allocate(bx(nx-1,ny,nz), by(nx,ny-1,nz), bz(nx,ny,nz-1), bp(nx,ny,nz))
      rc = aphid_build_rhs(ctx, int(mode, c_int), int(ns, c_int), sd,   &
                           si, sj, sk, skipped, bx, by, bz, bp)
      call check(rc, 'aphid_build_rhs')
      do s = 1, ns
        if (skipped(s) /= 0) write(6, '(i6,a)') s,                      &
          ' error: source index out of bounds or invalid... skipping.'
      end do

!     solve in place (the reference's rhs_* convention)
      allocate(hist(max_it + 1), htime(max_it + 1))
      rep%hist = c_loc(hist(1))
      rep%hist_time = c_loc(htime(1))
      rep%hist_cap = max_it + 1
      rc = aphid_qmr_solve(ctx, bx, by, bz, bp, c_null_ptr, c_null_ptr, &
                           c_null_ptr, c_null_ptr, target,              &
                           int(max_it, c_int), 0_c_int, rep)
      call check(rc, 'aphid_qmr_solve')

!     qmr.iter: iteration, residual norm, seconds since the solve started
!     (src/aphid_qmr.f90:381); the progress line every 10 iterations
!     (:376-379), printed after the solve since the iterations run on the
!     device
      open(unit=14, file=trim(wdir)//'/qmr.iter', status='replace')
      do i = 1, rep%hist_len
        write(14, '(i8,1x,es24.16,1x,es24.16)') i, hist(i), htime(i)
        if (mod(i, 10) == 0) write(6, 98) i, max_it, hist(i), htime(i)
      end do
      close(14)
   98 format('it: ',i6,'/',i6,'    r: ',e13.6,'    t:',e13.6)

      select case (rep%status)
        case (0); write(6, '(a,i8,a)') ' Solution converged in ',      &
                                         rep%iters, ' iterations'
        case (1); write(6, '(a)') ' ********** WARNING: maximum '//     &
                                  'iterations exceeded **********'
        case default; write(6, '(a)') ' ********** QMR breakdown '//    &
                                      '(non-finite recurrence) **********'
      end select
      write(6, '(a,es12.4)') '    Residual:', rep%rnorm
      write(6, '(a,es12.4)') ' True relative residual:', rep%true_rel_res
      write(6, '(a,f12.6)') ' Time spent in QMR [s]:', rep%seconds_qmr
      if (rep%iters > 0) write(6, '(a,es12.4)')                         &
        ' Seconds per iteration:', rep%seconds_qmr / rep%iters

      open(unit=15, file=trim(wdir)//'/aphid_solution.bin',             &
           access='stream', form='unformatted', status='replace')
      write(15) bx, by, bz, bp
      close(15)

!     post-processing as the reference's aphid_main (src/aphid.f90:110-112)
      call vtk_a_phi(trim(wdir)//'/aphid_a_phi.vtk', xn, yn, zn,        &
                     bx, by, bz, bp)
      allocate(fbx(nx,ny-1,nz-1), fby(nx-1,ny,nz-1), fbz(nx-1,ny-1,nz))
      allocate(fex(nx-1,ny,nz), fey(nx,ny-1,nz), fez(nx,ny,nz-1))
      rc = aphid_gen_fields(ctx, omega, fbx, fby, fbz, fex, fey, fez)
      call check(rc, 'aphid_gen_fields')
      allocate(k2x(nx-1,ny,nz), k2y(nx,ny-1,nz), k2z(nx,ny,nz-1),       &
               k2p(nx,ny,nz))
      rc = aphid_get_k2(ctx, k2x, k2y, k2z, k2p)
      call check(rc, 'aphid_get_k2')
      call vtk_fields(trim(wdir), xn, yn, zn, fbx, fby, fbz, fex, fey,   &
                      fez, k2x, k2y, k2z, omega)

      rc = aphid_ctx_destroy(ctx)

      contains

      subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(*), intent(in) :: what
        if (code < 0) then
          write(6, '(a)') ' >> ERROR in '//what//': '//                 &
                          trim(aphid_error_string())
          stop 1
        end if
      end subroutine check

      subroutine die(msg)
        character(*), intent(in) :: msg
        write(6, '(a)') ' >> ERROR '//msg
        stop 1
      end subroutine die

      end program aphid_gpu_serial
