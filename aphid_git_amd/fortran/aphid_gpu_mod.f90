!----------------------------------------------------------------------!
! aphid_gpu_mod.f90 -- iso_c_binding interfaces to libaphid_gpu.so     !
! (include/aphid_gpu.h).  A Fortran caller that today calls the        !
! reference's qmr / inprod_* / gen_diag (src/aphid.f90:102-104,        !
! src/aphid_qmr.f90:5) uses these instead.  Complex arrays are         !
! complex(c_double_complex), Fortran order, with the reference's       !
! shapes; optional arrays are passed as type(c_ptr) (c_loc or           !
! c_null_ptr).                                                          !
!----------------------------------------------------------------------!
      module aphid_gpu
      use iso_c_binding
      implicit none

      integer(c_int), parameter :: APHID_F_X0 = 1
      integer(c_int), parameter :: APHID_F_NO_TRUE_RES = 2
      integer(c_int), parameter :: APHID_F_MIXED = 4

      type, bind(C) :: aphid_qmr_report
        integer(c_int) :: iters = 0, status = 0
        real(c_double) :: rnorm = 0, rhsnorm = 0, true_rel_res = 0
        real(c_double) :: seconds_qmr = 0, seconds_total = 0
        type(c_ptr)    :: hist = c_null_ptr
        integer(c_int) :: hist_cap = 0, hist_len = 0
        integer(c_int) :: outer_iters = 0, reserved = 0
        type(c_ptr)    :: hist_time = c_null_ptr
      end type aphid_qmr_report

      type, bind(C) :: aphid_sys_report
        integer(c_int) :: iters, status
        real(c_double) :: rnorm, rhsnorm, true_rel_res, r0_rel
      end type aphid_sys_report

      interface
        integer(c_int) function aphid_ctx_create(ctx, device)          &
                                bind(C, name='aphid_ctx_create')
          import :: c_int, c_ptr
          type(c_ptr), intent(out) :: ctx
          integer(c_int), value :: device
        end function

        integer(c_int) function aphid_ctx_destroy(ctx)                 &
                                bind(C, name='aphid_ctx_destroy')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function

        type(c_ptr) function aphid_last_error()                        &
                             bind(C, name='aphid_last_error')
          import :: c_ptr
        end function

        integer(c_int) function aphid_set_mesh(ctx, nx, ny, nz,        &
                         xn, yn, zn) bind(C, name='aphid_set_mesh')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: nx, ny, nz
          real(c_double), intent(in) :: xn(*), yn(*), zn(*)
        end function

        integer(c_int) function aphid_set_cells(ctx, model, omega,     &
                         a, b, c, d) bind(C, name='aphid_set_cells')
          import :: c_int, c_ptr, c_double, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), value :: model
          real(c_double), value :: omega
          type(c_ptr), value :: a, b, c, d
        end function

        integer(c_int) function aphid_set_sigma(ctx, sig, omega)       &
                                bind(C, name='aphid_set_sigma')
          import :: c_int, c_ptr, c_double, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(in) :: sig(*)
          real(c_double), value :: omega
        end function

        integer(c_int) function aphid_set_k2(ctx, kx, ky, kz, kp)      &
                                bind(C, name='aphid_set_k2')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(in) :: kx(*), ky(*), kz(*), kp(*)
        end function

        integer(c_int) function aphid_set_diag(ctx, dx, dy, dz, dp)    &
                                bind(C, name='aphid_set_diag')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(in) :: dx(*), dy(*), dz(*), dp(*)
        end function

        integer(c_int) function aphid_get_k2(ctx, kx, ky, kz, kp)      &
                                bind(C, name='aphid_get_k2')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(out) :: kx(*), ky(*), kz(*), kp(*)
        end function

        integer(c_int) function aphid_get_sigma(ctx, sig)              &
                                bind(C, name='aphid_get_sigma')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(out) :: sig(*)
        end function

        integer(c_int) function aphid_gen_fields(ctx, omega, bx, by,   &
                                bz, ex, ey, ez)                        &
                                bind(C, name='aphid_gen_fields')
          import :: c_int, c_ptr, c_double, c_double_complex
          type(c_ptr), value :: ctx
          real(c_double), value :: omega
          complex(c_double_complex), intent(out) :: bx(*), by(*), bz(*)
          complex(c_double_complex), intent(out) :: ex(*), ey(*), ez(*)
        end function

        integer(c_int) function aphid_get_diag(ctx, dx, dy, dz, dp)    &
                                bind(C, name='aphid_get_diag')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(out) :: dx(*), dy(*), dz(*), dp(*)
        end function

        integer(c_int) function aphid_apply(ctx, ix, iy, iz, ip,       &
                         ox, oy, oz, op) bind(C, name='aphid_apply')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(in)  :: ix(*), iy(*), iz(*), ip(*)
          complex(c_double_complex), intent(out) :: ox(*), oy(*), oz(*), op(*)
        end function

        integer(c_int) function aphid_build_rhs(ctx, mode, nsrc, dir,  &
                         ip, jp, kp, skipped, bx, by, bz, bp)           &
                         bind(C, name='aphid_build_rhs')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          integer(c_int), value :: mode, nsrc
          integer(c_int), intent(in) :: dir(*), ip(*), jp(*), kp(*)
          integer(c_int), intent(out) :: skipped(*)
          complex(c_double_complex), intent(out) :: bx(*), by(*), bz(*), bp(*)
        end function

        integer(c_int) function aphid_qmr_solve(ctx, xx, xy, xz, xp,   &
                         x0x, x0y, x0z, x0p, tol, max_it, flags, rep)   &
                         bind(C, name='aphid_qmr_solve')
          import :: c_int, c_ptr, c_double, c_double_complex,          &
                    aphid_qmr_report
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(inout) :: xx(*), xy(*), xz(*), xp(*)
          type(c_ptr), value :: x0x, x0y, x0z, x0p
          real(c_double), value :: tol
          integer(c_int), value :: max_it, flags
          type(aphid_qmr_report), intent(inout) :: rep
        end function

        integer(c_int) function aphid_set_reproducible(ctx, on)       &
                         bind(C, name='aphid_set_reproducible')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: on
        end function

        integer(c_int) function aphid_load_rhs(ctx, bx, by, bz, bp,    &
                         x0x, x0y, x0z, x0p) bind(C, name='aphid_load_rhs')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(in) :: bx(*), by(*), bz(*), bp(*)
          type(c_ptr), value :: x0x, x0y, x0z, x0p
        end function

        integer(c_int) function aphid_solve_resident(ctx, tol, max_it, &
                         flags, rep) bind(C, name='aphid_solve_resident')
          import :: c_int, c_ptr, c_double, aphid_qmr_report
          type(c_ptr), value :: ctx
          real(c_double), value :: tol
          integer(c_int), value :: max_it, flags
          type(aphid_qmr_report), intent(inout) :: rep
        end function

        integer(c_int) function aphid_fetch_solution(ctx, xx, xy, xz,  &
                         xp) bind(C, name='aphid_fetch_solution')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: ctx
          complex(c_double_complex), intent(out) :: xx(*), xy(*), xz(*), xp(*)
        end function

        integer(c_int) function aphid_batch_create(batch, device)      &
                                bind(C, name='aphid_batch_create')
          import :: c_int, c_ptr
          type(c_ptr), intent(out) :: batch
          integer(c_int), value :: device
        end function

        integer(c_int) function aphid_batch_destroy(batch)             &
                                bind(C, name='aphid_batch_destroy')
          import :: c_int, c_ptr
          type(c_ptr), value :: batch
        end function

        integer(c_int) function aphid_batch_add_operator(batch, nx, ny,&
                                nz, xn, yn, zn, model, omega, a, b, c, d)&
                                bind(C, name='aphid_batch_add_operator')
          import :: c_int, c_ptr, c_double, c_char
          type(c_ptr), value :: batch
          integer(c_int), value :: nx, ny, nz
          real(c_double), intent(in) :: xn(*), yn(*), zn(*)
          character(kind=c_char), value :: model
          real(c_double), value :: omega
          type(c_ptr), value :: a, b, c, d
        end function

        integer(c_int) function aphid_batch_add_system(batch, op, bx,  &
                                by, bz, bp, x0x, x0y, x0z, x0p)        &
                                bind(C, name='aphid_batch_add_system')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: batch
          integer(c_int), value :: op
          complex(c_double_complex), intent(in) :: bx(*), by(*), bz(*), bp(*)
          type(c_ptr), value :: x0x, x0y, x0z, x0p
        end function

        integer(c_int) function aphid_batch_solve(batch, tol, max_it,  &
                                flags, reports, seconds)               &
                                bind(C, name='aphid_batch_solve')
          import :: c_int, c_ptr, c_double, aphid_sys_report
          type(c_ptr), value :: batch
          real(c_double), value :: tol
          integer(c_int), value :: max_it, flags
          type(aphid_sys_report), intent(out) :: reports(*)
          real(c_double), intent(out) :: seconds
        end function

        integer(c_int) function aphid_batch_set_sweep_guess(batch,    &
                                sys, nfrom, from)                      &
                                bind(C, name='aphid_batch_set_sweep_guess')
          import :: c_int, c_ptr
          type(c_ptr), value :: batch
          integer(c_int), value :: sys, nfrom
          integer(c_int), intent(in) :: from(*)
        end function

        integer(c_int) function aphid_batch_set_history(batch, cap)    &
                                bind(C, name='aphid_batch_set_history')
          import :: c_int, c_ptr
          type(c_ptr), value :: batch
          integer(c_int), value :: cap
        end function

        integer(c_int) function aphid_batch_get_history(batch, sys,    &
                                hist, time, cap, len)                  &
                                bind(C, name='aphid_batch_get_history')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: batch
          integer(c_int), value :: sys, cap
          real(c_double), intent(out) :: hist(*), time(*)
          integer(c_int), intent(out) :: len
        end function

        integer(c_int) function aphid_batch_get_solution(batch, sys,   &
                                xx, xy, xz, xp)                        &
                                bind(C, name='aphid_batch_get_solution')
          import :: c_int, c_ptr, c_double_complex
          type(c_ptr), value :: batch
          integer(c_int), value :: sys
          complex(c_double_complex), intent(out) :: xx(*), xy(*), xz(*), xp(*)
        end function
      end interface

      contains

!     the library's last error message as a Fortran string
      function aphid_error_string() result(msg)
        character(len=256) :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: s(:)
        integer :: n
        msg = ' '
        p = aphid_last_error()
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [256])
        do n = 1, 256
          if (s(n) == c_null_char) exit
          msg(n:n) = s(n)
        end do
      end function aphid_error_string

      end module aphid_gpu
