!----------------------------------------------------------------------!
! ref_harness.f90 -- drives the REFERENCE's own routines (compiled from  !
! their sources under /root/reference/src by oracle/Makefile) to write  !
! full-precision golden vectors and to time the reference QMR.          !
!                                                                      !
! TEST INFRASTRUCTURE.  Run from a case directory holding aphid.input,  !
! aphid.mesh (and aphid.source for mode 5).  Commands:                  !
!   ref_harness setup        -> ref_setup.bin                          !
!   ref_harness apply N      reads ref_apply_in.bin (N 4-component     !
!                            vectors), writes ref_apply_out.bin        !
!   ref_harness qmr1|qmr4    -> ref_qmr_sol.bin (+ the reference's own  !
!                            qmr.iter); qmr1 = serial qmr, qmr4 =      !
!                            4-thread qmr_mthread                      !
!   ref_harness fields       serial qmr, then gen_fields and the VTK    !
!                            writers: ref_qmr_sol.bin, ref_fields.bin   !
!                            (Bx,By,Bz,Ex,Ey,Ez), sigma.vtk, Bx..Ez.vtk !
!                            and ref_a_phi.vtk (dump_a_phi's unit 11)   !
! Binary files are raw stream records in the order written below.      !
!----------------------------------------------------------------------!
      program ref_harness
      use general
      use mesh
      implicit none
      character(32) :: cmd, arg
      integer :: nvec, v, c0, c1, crate
      complex(dp), allocatable :: ax(:,:,:), ay(:,:,:), az(:,:,:), ap(:,:,:)
      complex(dp), allocatable :: ox(:,:,:), oy(:,:,:), oz(:,:,:), op(:,:,:)

      call get_command_argument(1, cmd)
      pwd = './'
!     the reference's QMR times itself with mpi_wtime, which MPICH only
!     serves after MPI_Init (singleton init, no launcher needed)
      call mpi_init(ierr)

      call read_input
      call read_mesh
      call gen_rhs
      call gen_diag
      call gen_inprod_x_coefs
      call gen_inprod_y_coefs
      call gen_inprod_z_coefs
      call gen_inprod_p_coefs

      if (trim(cmd) .eq. 'setup') then
        open(unit=21, file='ref_setup.bin', access='stream',            &
             form='unformatted', status='replace')
        write(21) nx, ny, nz
        write(21) dx, dy, dz, dxc, dyc, dzc
        write(21) sig
        write(21) k2_xgrid, k2_ygrid, k2_zgrid, k2_pgrid
        write(21) diag_x, diag_y, diag_z, diag_p
        write(21) rhs_x, rhs_y, rhs_z, rhs_p
        close(21)

      else if (trim(cmd) .eq. 'apply') then
        call get_command_argument(2, arg)
        read(arg, *) nvec
        allocate(ax(nx-1,ny,nz), ay(nx,ny-1,nz), az(nx,ny,nz-1),        &
                 ap(nx,ny,nz))
        allocate(ox(nx-1,ny,nz), oy(nx,ny-1,nz), oz(nx,ny,nz-1),        &
                 op(nx,ny,nz))
        open(unit=22, file='ref_apply_in.bin', access='stream',         &
             form='unformatted', status='old')
        open(unit=23, file='ref_apply_out.bin', access='stream',        &
             form='unformatted', status='replace')
        do v = 1, nvec
          read(22) ax, ay, az, ap
          call inprod_x(ax, ay, az, ap, ox)
          call inprod_y(ax, ay, az, ap, oy)
          call inprod_z(ax, ay, az, ap, oz)
          call inprod_p(ax, ay, az, ap, op)
          write(23) ox, oy, oz, op
        end do
        close(22)
        close(23)

      else if (trim(cmd) .eq. 'qmr1' .or. trim(cmd) .eq. 'qmr4') then
        call system_clock(c0, crate)
        if (trim(cmd) .eq. 'qmr1') call qmr
        if (trim(cmd) .eq. 'qmr4') call qmr_mthread
        call system_clock(c1)
        write(6,'(a,es24.16)') ' HARNESS qmr_wall_seconds ',              &
                               dble(c1-c0)/dble(crate)
        open(unit=24, file='ref_qmr_sol.bin', access='stream',          &
             form='unformatted', status='replace')
        write(24) rhs_x, rhs_y, rhs_z, rhs_p
        close(24)
      else if (trim(cmd) .eq. 'fields') then
        call dump_sigma
        call qmr
        open(unit=24, file='ref_qmr_sol.bin', access='stream',          &
             form='unformatted', status='replace')
        write(24) rhs_x, rhs_y, rhs_z, rhs_p
        close(24)
!       dump_a_phi writes to unit 11 without opening it
        open(unit=11, file='ref_a_phi.vtk', status='replace')
        call dump_a_phi
        call gen_fields
        call dump_fields
        open(unit=25, file='ref_fields.bin', access='stream',           &
             form='unformatted', status='replace')
        write(25) Bx, By, Bz, Ex, Ey, Ez
        close(25)
      else
        write(6,*) 'usage: ref_harness setup | apply N | qmr1 | qmr4 | fields'
        call mpi_finalize(ierr)
        stop 2
      end if
      call mpi_finalize(ierr)
      end program ref_harness
