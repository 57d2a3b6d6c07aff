! End to end through the Fortran binding on the GPU: a 40x32x28 stretched
! grid with one rotated ellipsoid (pp_geom_build, triangulated), pp_create,
! pp_set_geometry, pp_solve of an integer-pattern right-hand side to 1e-10,
! and the nodes, iblank and solution written to the file named on the
! command line (stream access, the whole (0:nx+2,0:ny+2,0:nz+2) arrays) for
! the test to repeat the same solve through the ctypes binding.
program solve_e2e
  use iso_c_binding
  use pp_hip_mod
  implicit none
  integer, parameter :: nxc = 40, nyc = 32, nzc = 28
  type(pp_grid_desc) :: g
  type(pp_geometry) :: geom
  type(pp_solve_opts) :: o
  type(pp_solve_stats) :: st
  type(pp_body_desc) :: body(1)
  type(c_ptr) :: ctx, geo
  integer(c_int32_t) :: rc, ng, nb, nf
  real(c_double), target :: xn(nxc + 1), yn(nyc + 1), zn(nzc + 1)
  integer(1), allocatable, target :: ib(:,:,:)
  integer(c_int32_t), allocatable, target :: gi(:,:)
  real(c_double), allocatable, target :: coeff(:,:), probe(:), p(:,:,:), rhs(:,:,:)
  integer :: i, j, k, u
  character(len=512) :: path
  call get_command_argument(1, path)
  ! nodes: x uniform on [0, 1.25], y and z quadratic clustering on [0, 1]
  xn = [(1.25d0 * i / nxc, i = 0, nxc)]
  yn = [(dble(i) / nyc, i = 0, nyc)]
  zn = [(dble(i) / nzc, i = 0, nzc)]
  yn = yn * yn * (3.0d0 - 2.0d0 * yn)
  zn = zn * zn * (3.0d0 - 2.0d0 * zn)
  g%nxc = nxc; g%nyc = nyc; g%nzc = nzc
  g%x = c_loc(xn); g%y = c_loc(yn); g%z = c_loc(zn)
  g%bc = PP_BC_NEUMANN
  g%device = 0
  body(1)%shape = PP_BODY_ELLIPSOID
  body(1)%center = [0.6d0, 0.48d0, 0.52d0]
  body(1)%semi = [0.22d0, 0.16d0, 0.18d0]
  body(1)%alpha_deg = 25.0d0
  body(1)%n_theta = 40
  rc = pp_geom_build(geo, g, 1, body, PP_GEOM_TRIANGULATED, 2.0d0, c_null_ptr)
  if (rc /= PP_OK) call fail('pp_geom_build', rc)
  rc = pp_geom_counts(geo, ng, nb, nf)
  allocate(ib(0:nxc + 2, 0:nyc + 2, 0:nzc + 2), gi(ng, 6), coeff(8, ng), probe(ng))
  rc = pp_geom_get(geo, c_loc(ib), c_null_ptr, c_loc(gi(1, 1)), c_loc(gi(1, 2)), c_loc(gi(1, 3)), &
       c_loc(gi(1, 4)), c_loc(gi(1, 5)), c_loc(gi(1, 6)), c_loc(coeff), c_loc(probe), c_null_ptr, c_null_ptr, &
       c_null_ptr)
  if (rc /= PP_OK) call fail('pp_geom_get', rc)
  rc = pp_geom_destroy(geo)
  rc = pp_create(ctx, g)
  if (rc /= PP_OK) call fail('pp_create', rc)
  geom%iblank = c_loc(ib)
  geom%fresh_cell = c_null_ptr
  geom%n_ghost = ng
  geom%i_ghost = c_loc(gi(1, 1)); geom%j_ghost = c_loc(gi(1, 2)); geom%k_ghost = c_loc(gi(1, 3))
  geom%i_cell = c_loc(gi(1, 4)); geom%j_cell = c_loc(gi(1, 5)); geom%k_cell = c_loc(gi(1, 6))
  geom%coeff = c_loc(coeff)
  geom%probe_length = c_loc(probe)
  rc = pp_set_geometry(ctx, geom)
  if (rc /= PP_OK) call fail('pp_set_geometry', rc)
  allocate(p(0:nxc + 2, 0:nyc + 2, 0:nzc + 2), rhs(0:nxc + 2, 0:nyc + 2, 0:nzc + 2))
  p = 0.0d0
  rhs = 0.0d0
  do k = 1, nzc
     do j = 1, nyc
        do i = 1, nxc
           if (ib(i, j, k) == 0) rhs(i, j, k) = dble(mod(7 * i + 13 * j + 17 * k, 29) - 14) / 14.0d0
        end do
     end do
  end do
  call pp_default_opts(o)
  o%rel_tol = 1.0d-10
  o%max_outer = 100
  rc = pp_solve(ctx, c_loc(p), c_loc(rhs), c_null_ptr, o, st)
  if (rc /= PP_OK) call fail('pp_solve', rc)
  print '(A,I0,1X,I0,1X,ES24.16,1X,I0)', 'solve ', st%status, st%outer_iterations, st%rel_l2, ng
  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
  write(u) xn, yn, zn, ib, p
  close(u)
  rc = pp_destroy(ctx)
  print '(A,I0)', 'destroy ', rc
contains
  subroutine fail(what, code)
    character(len=*), intent(in) :: what
    integer(c_int32_t), intent(in) :: code
    print '(A,1X,A,1X,I0,1X,A)', 'FAIL', what, code, pp_error_text()
    stop 1
  end subroutine fail
end program solve_e2e
