!----------------------------------------------------------------------!
! aphid_vtk.f90 -- the reference's VTK output (rectilinear grids, ASCII, !
! values as e14.7, up to 1000 per record) for the GPU driver:          !
!   vtk_sigma    dump_sigma   (src/aphid.f90:1803-1838)                 !
!   vtk_a_phi    dump_a_phi   (:1731-1796), written to aphid_a_phi.vtk  !
!                (the reference writes it to unit 11 without opening    !
!                it, i.e. to the compiler's default file for unit 11)   !
!   vtk_fields   dump_fields  (:1844-2166): Bx..Ez.vtk and EM_grid.vtk  !
! Header lines, count widths and the node averages of EM_grid.vtk       !
! follow the reference line by line, including its By average (whose    !
! third term repeats the second, :2138-2146), so the files compare      !
! value for value.                                                      !
!----------------------------------------------------------------------!
      module aphid_vtk
      implicit none
      private
      public :: vtk_sigma, vtk_a_phi, vtk_fields

      integer, parameter :: dp = 8
      real(dp), parameter :: PI = 3.141592653589793d0
      real(dp), parameter :: MU0 = PI * 4.d-7
      complex(dp), parameter :: II = (0.d0, 1.d0)
      integer, parameter :: u = 31

      contains

      subroutine head(path, n1, n2, n3, xc, yc, zc)
        character(*), intent(in) :: path
        integer, intent(in) :: n1, n2, n3
        real(dp), intent(in) :: xc(:), yc(:), zc(:)
        open(unit=u, file=path, status='replace')
        write(u,'(a)') '# vtk DataFile Version 3.0'
        write(u,'(a)') 'vtk output'
        write(u,'(a)') 'ASCII'
        write(u,'(a)') 'DATASET RECTILINEAR_GRID'
        write(u,'(a,3(i3,1x))') 'DIMENSIONS ',n1,n2,n3
        write(u,'(a,i8,a)') 'X_COORDINATES ',n1,' float'
        write(u,1) xc
        write(u,'(a,i8,a)') 'Y_COORDINATES ',n2,' float'
        write(u,1) yc
        write(u,'(a,i8,a)') 'Z_COORDINATES ',n3,' float'
        write(u,1) zc
 1      format(1000(e14.7,1x))
      end subroutine head

      subroutine scal(name, kind, v)
        character(*), intent(in) :: name, kind
        real(dp), intent(in) :: v(:,:,:)
        write(u,'(a)') 'SCALARS '//name//' '//kind//' 1'
        write(u,'(a)') 'LOOKUP_TABLE default'
        write(u,1) v
 1      format(1000(e14.7,1x))
      end subroutine scal

      function mid(c) result(m)
        real(dp), intent(in) :: c(:)
        real(dp) :: m(size(c)-1)
        m = 0.5d0*( c(1:size(c)-1) + c(2:size(c)) )
      end function mid

      subroutine vtk_sigma(path, xn, yn, zn, sig)
        character(*), intent(in) :: path
        real(dp), intent(in) :: xn(:), yn(:), zn(:)
        complex(dp), intent(in) :: sig(:,:,:)
        integer :: nx, ny, nz
        nx = size(xn); ny = size(yn); nz = size(zn)
        call head(path, nx, ny, nz, xn, yn, zn)
        write(u,'(a,i14,a)') 'CELL_DATA ' ,(nx-1)*(ny-1)*(nz-1)
        call scal('sig_real', 'float', real(sig))
        call scal('sig_imag', 'float', imag(sig))
        close(u)
      end subroutine vtk_sigma

      subroutine vtk_a_phi(path, xn, yn, zn, ax, ay, az, ap)
        character(*), intent(in) :: path
        real(dp), intent(in) :: xn(:), yn(:), zn(:)
        complex(dp), intent(in) :: ax(:,:,:), ay(:,:,:), az(:,:,:), ap(:,:,:)
        integer :: nx, ny, nz
        nx = size(xn); ny = size(yn); nz = size(zn)
        call head(path, nx-2, ny-2, nz-2, xn(2:nx-1), yn(2:ny-1), zn(2:nz-1))
        write(u,'(a,i8,a)') 'POINT_DATA ' ,(nx-2)*(ny-2)*(nz-2)
        call scal('Ax_real', 'double', 0.5 * real(ax(1:nx-2,2:ny-1,2:nz-1) +  &
                                              ax(2:nx-1,2:ny-1,2:nz-1)))
        call scal('Ax_imag', 'double', 0.5 * imag(ax(1:nx-2,2:ny-1,2:nz-1) +  &
                                              ax(2:nx-1,2:ny-1,2:nz-1)))
        call scal('Ay_real', 'double', 0.5 * real(ay(2:nx-1,1:ny-2,2:nz-1) +  &
                                              ay(2:nx-1,2:ny-1,2:nz-1)))
        call scal('Ay_imag', 'double', 0.5 * imag(ay(2:nx-1,1:ny-2,2:nz-1) +  &
                                              ay(2:nx-1,2:ny-1,2:nz-1)))
        call scal('Az_real', 'double', 0.5 * real(az(2:nx-1,2:ny-1,1:nz-2) +  &
                                              az(2:nx-1,2:ny-1,2:nz-1)))
        call scal('Az_imag', 'double', 0.5 * imag(az(2:nx-1,2:ny-1,1:nz-2) +  &
                                              az(2:nx-1,2:ny-1,2:nz-1)))
        call scal('phi_real', 'double', real(ap(2:nx-1,2:ny-1,2:nz-1)))
        call scal('phi_imag', 'double', imag(ap(2:nx-1,2:ny-1,2:nz-1)))
        close(u)
      end subroutine vtk_a_phi

      subroutine vtk_fields(dir, xn, yn, zn, bx, by, bz, ex, ey, ez,      &
                            k2x, k2y, k2z, omega)
        character(*), intent(in) :: dir
        real(dp), intent(in) :: xn(:), yn(:), zn(:), omega
        complex(dp), intent(in) :: bx(:,:,:), by(:,:,:), bz(:,:,:)
        complex(dp), intent(in) :: ex(:,:,:), ey(:,:,:), ez(:,:,:)
        complex(dp), intent(in) :: k2x(:,:,:), k2y(:,:,:), k2z(:,:,:)
        integer :: nx, ny, nz
        nx = size(xn); ny = size(yn); nz = size(zn)

        call head(dir//'/Bx.vtk', nx, ny-1, nz-1, xn, mid(yn), mid(zn))
        write(u,'(a,i14,a)') 'POINT_DATA ' ,nx*(ny-1)*(nz-1)
        call scal('Bx_real', 'double', real(bx))
        call scal('Bx_imag', 'double', imag(bx))
        close(u)
        call head(dir//'/By.vtk', nx-1, ny, nz-1, mid(xn), yn, mid(zn))
        write(u,'(a,i14,a)') 'POINT_DATA ' ,(nx-1)*ny*(nz-1)
        call scal('By_real', 'double', real(by))
        call scal('By_imag', 'double', imag(by))
        close(u)
        call head(dir//'/Bz.vtk', nx-1, ny-1, nz, mid(xn), mid(yn), zn)
        write(u,'(a,i14,a)') 'POINT_DATA ' ,(nx-1)*(ny-1)*nz
        call scal('Bz_real', 'double', real(bz))
        call scal('Bz_imag', 'double', imag(bz))
        close(u)
        call head(dir//'/Ex.vtk', nx-1, ny, nz, mid(xn), yn, zn)
        write(u,'(a,i14,a)') 'POINT_DATA ' ,(nx-1)*ny*nz
        call scal('Ex_real', 'double', real(ex))
        call scal('Ex_imag', 'double', imag(ex))
        close(u)
        call head(dir//'/Ey.vtk', nx, ny-1, nz, xn, mid(yn), zn)
        write(u,'(a,i8,a)') 'POINT_DATA ' ,nx*(ny-1)*nz
        call scal('Ey_real', 'double', real(ey))
        call scal('Ey_imag', 'double', imag(ey))
        close(u)
        call head(dir//'/Ez.vtk', nx, ny, nz-1, xn, yn, mid(zn))
        write(u,'(a,i8,a)') 'POINT_DATA ' ,nx*ny*(nz-1)
        call scal('Ez_real', 'double', real(ez))
        call scal('Ez_imag', 'double', imag(ez))
        close(u)

!       everything averaged to the interior nodes
        call head(dir//'/EM_grid.vtk', nx-2, ny-2, nz-2, xn(2:nx-1),     &
                  yn(2:ny-1), zn(2:nz-1))
        write(u,'(a,i8,a)') 'POINT_DATA ' ,(nx-2)*(ny-2)*(nz-2)
        call scal('Ex_grid_real', 'double', real( ( ex(1:nx-2,2:ny-1,2:nz-1) + &
                                                    ex(2:nx-1,2:ny-1,2:nz-1) )/2.))
        call scal('Ex_grid_imag', 'double', imag( ( ex(1:nx-2,2:ny-1,2:nz-1) + &
                                                    ex(2:nx-1,2:ny-1,2:nz-1) )/2.))
        call scal('Ey_grid_real', 'double', real( ( ey(2:nx-1,1:ny-2,2:nz-1) + &
                                                    ey(2:nx-1,2:ny-1,2:nz-1) )/2.))
        call scal('Ey_grid_imag', 'double', imag( ( ey(2:nx-1,1:ny-2,2:nz-1) + &
                                                    ey(2:nx-1,2:ny-1,2:nz-1) )/2.))
        call scal('Ez_grid_real', 'double', real( ( ez(2:nx-1,2:ny-1,1:nz-2) + &
                                                    ez(2:nx-1,2:ny-1,2:nz-1) )/2.))
        call scal('Ez_grid_imag', 'double', imag( ( ez(2:nx-1,2:ny-1,1:nz-2) + &
                                                    ez(2:nx-1,2:ny-1,2:nz-1) )/2.))
        call scal('Jx_grid_real', 'double', real( (                          &
              ex(1:nx-2,2:ny-1,2:nz-1)*k2x(1:nx-2,2:ny-1,2:nz-1) +            &
              ex(2:nx-1,2:ny-1,2:nz-1)*k2x(2:nx-1,2:ny-1,2:nz-1) ) /          &
                                                (2. * II * omega * MU0) ))
        call scal('Jx_grid_imag', 'double', imag( (                          &
              ex(1:nx-2,2:ny-1,2:nz-1)*k2x(1:nx-2,2:ny-1,2:nz-1) +            &
              ex(2:nx-1,2:ny-1,2:nz-1)*k2x(2:nx-1,2:ny-1,2:nz-1) ) /          &
                                                (2. * II * omega * MU0) ))
        call scal('Jy_grid_real', 'double', real( (                          &
              ey(2:nx-1,1:ny-2,2:nz-1)*k2y(2:nx-1,1:ny-2,2:nz-1) +            &
              ey(2:nx-1,2:ny-1,2:nz-1)*k2y(2:nx-1,2:ny-1,2:nz-1) ) /          &
                                                (2. * II * omega * MU0) ))
        call scal('Jy_grid_imag', 'double', imag( (                          &
              ey(2:nx-1,1:ny-2,2:nz-1)*k2y(2:nx-1,1:ny-2,2:nz-1) +            &
              ey(2:nx-1,2:ny-1,2:nz-1)*k2y(2:nx-1,2:ny-1,2:nz-1) ) /          &
                                                (2. * II * omega * MU0) ))
        call scal('Jz_grid_real', 'double', real( (                          &
              ez(2:nx-1,2:ny-1,2:nz-1)*k2z(2:nx-1,2:ny-1,2:nz-1) +            &
              ez(2:nx-1,2:ny-1,1:nz-2)*k2z(2:nx-1,2:ny-1,1:nz-2) ) /          &
                                                (2. * II * omega * MU0) ))
        call scal('Jz_grid_imag', 'double', imag( (                          &
              ez(2:nx-1,2:ny-1,2:nz-1)*k2z(2:nx-1,2:ny-1,2:nz-1) +            &
              ez(2:nx-1,2:ny-1,1:nz-2)*k2z(2:nx-1,2:ny-1,1:nz-2) ) /          &
                                                (2. * II * omega * MU0) ))
        call scal('Bx_grid_real', 'double', real( (                          &
                 bx(2:nx-1,1:ny-2,1:nz-2) + bx(2:nx-1,2:ny-1,1:nz-2) +       &
                 bx(2:nx-1,1:ny-2,2:nz-1) + bx(2:nx-1,2:ny-1,2:nz-1) )/4.))
        call scal('Bx_grid_imag', 'double', imag( (                          &
                 bx(2:nx-1,1:ny-2,1:nz-2) + bx(2:nx-1,2:ny-1,1:nz-2) +       &
                 bx(2:nx-1,1:ny-2,2:nz-1) + bx(2:nx-1,2:ny-1,2:nz-1) )/4.))
        call scal('By_grid_real', 'double', real( (                          &
                 by(1:nx-2,2:ny-1,1:nz-2) + by(2:nx-1,2:ny-1,1:nz-2) +       &
                 by(2:nx-1,2:ny-1,1:nz-2) + by(2:nx-1,2:ny-1,2:nz-1) )/4.))
        call scal('By_grid_imag', 'double', imag( (                          &
                 by(1:nx-2,2:ny-1,1:nz-2) + by(2:nx-1,2:ny-1,1:nz-2) +       &
                 by(2:nx-1,2:ny-1,1:nz-2) + by(2:nx-1,2:ny-1,2:nz-1) )/4.))
        call scal('Bz_grid_real', 'double', real( (                          &
                 bz(1:nx-2,1:ny-2,2:nz-1) + bz(1:nx-2,2:ny-1,2:nz-1) +       &
                 bz(2:nx-1,1:ny-2,2:nz-1) + bz(2:nx-1,2:ny-1,2:nz-1) )/4.))
        call scal('Bz_grid_imag', 'double', imag( (                          &
                 bz(1:nx-2,1:ny-2,2:nz-1) + bz(1:nx-2,2:ny-1,2:nz-1) +       &
                 bz(2:nx-1,1:ny-2,2:nz-1) + bz(2:nx-1,2:ny-1,2:nz-1) )/4.))
        close(u)
      end subroutine vtk_fields

      end module aphid_vtk
