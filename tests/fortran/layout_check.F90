! Links the Fortran binding against the library and prints what the C side
! wrote into the BIND(C) types (no GPU needed): pp_default_opts, the status
! and error text of pp_create with a bad grid, the type sizes, and a host-only
! triangulated-ellipsoid geometry built through pp_geom_build on 24^3 cells
! (25 nodes per direction, fields (0:26)^3)
! (counts and checksums of what pp_geom_get and pp_geom_ghost_elements return).
program layout_check
  use iso_c_binding
  use pp_hip_mod
  implicit none
  type(pp_solve_opts) :: o
  type(pp_grid_desc) :: g
  type(pp_solve_stats) :: st
  type(pp_ad_stats) :: ast
  type(pp_surface_desc) :: srf
  type(c_ptr) :: ctx
  integer(c_int32_t) :: rc
  real(c_double), target :: xs(3), xn(25)
  type(pp_body_desc) :: body(1)
  type(c_ptr) :: geo
  integer(c_int32_t) :: ng, nb, nf, i
  integer(1), allocatable, target :: ib(:,:,:)
  integer(c_int32_t), allocatable, target :: gi(:,:), gb(:), ge(:)
  real(c_double), allocatable, target :: coeff(:,:), probe(:), coeffD(:,:)
  call pp_default_opts(o)
  print '(A,I0,1X,ES12.5,1X,ES12.5,7(1X,I0),1X,F6.3,5(1X,I0))', 'opts ', o%precision, o%rel_tol, o%abs_linf_tol, &
       o%max_outer, o%cycles_per_outer, o%nu_pre, o%nu_post, o%mean_remove, o%project_rhs, o%use_graph, &
       o%omega, o%profile, o%krylov, o%nu_coarse, o%smoother, 0
  xs = [0.0d0, 0.5d0, 1.0d0]
  g%nxc = 2; g%nyc = 2; g%nzc = 2
  g%x = c_loc(xs); g%y = c_loc(xs); g%z = c_loc(xs)
  g%bc = PP_BC_NEUMANN
  g%device = 0
  rc = pp_create(ctx, g)
  print '(A,I0)', 'create ', rc
  print '(A,A)', 'error ', pp_error_text()
  print '(A,I0,1X,I0,1X,I0,1X,I0)', 'sizes ', storage_size(o) / 8, storage_size(g) / 8, storage_size(st) / 8, &
       storage_size(pp_geometry(c_null_ptr, c_null_ptr, 0, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
       c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr)) / 8
  print '(A,I0)', 'bodysize ', storage_size(body(1)) / 8
  print '(A,I0)', 'adstatsize ', storage_size(ast) / 8
  print '(A,I0)', 'surfsize ', storage_size(srf) / 8
  xn = [(1.2d0 * i / 24.0d0, i = 0, 24)]
  g%nxc = 24; g%nyc = 24; g%nzc = 24
  g%x = c_loc(xn); g%y = c_loc(xn); g%z = c_loc(xn)
  body(1)%shape = PP_BODY_ELLIPSOID
  body(1)%center = [0.61d0, 0.58d0, 0.6d0]
  body(1)%semi = [0.3d0, 0.2d0, 0.22d0]
  body(1)%alpha_deg = 20.0d0
  body(1)%n_theta = 40
  rc = pp_geom_build(geo, g, 1, body, PP_GEOM_TRIANGULATED, 2.0d0, c_null_ptr)
  print '(A,I0)', 'geom ', rc
  rc = pp_geom_counts(geo, ng, nb, nf)
  print '(A,I0,1X,I0,1X,I0,1X,I0)', 'counts ', rc, ng, nb, nf
  allocate(ib(0:26, 0:26, 0:26), gi(ng, 6), gb(ng), ge(ng), coeff(8, ng), probe(ng))
  rc = pp_geom_get(geo, c_loc(ib), c_null_ptr, c_loc(gi(1, 1)), c_loc(gi(1, 2)), c_loc(gi(1, 3)), &
       c_loc(gi(1, 4)), c_loc(gi(1, 5)), c_loc(gi(1, 6)), c_loc(coeff), c_loc(probe), c_null_ptr, c_null_ptr, &
       c_null_ptr)
  print '(A,I0,1X,I0,6(1X,I0),1X,ES24.16,1X,ES24.16)', 'get ', rc, sum(int(ib)), sum(gi(:, 1)), sum(gi(:, 2)), &
       sum(gi(:, 3)), sum(gi(:, 4)), sum(gi(:, 5)), sum(gi(:, 6)), sum(coeff), sum(probe)
  rc = pp_geom_ghost_elements(geo, c_loc(gb), c_loc(ge))
  print '(A,I0,1X,I0,1X,I0)', 'elems ', rc, sum(gb), sum(ge)
  allocate(coeffD(8, ng))
  rc = pp_geom_dirichlet(geo, c_loc(coeffD))
  print '(A,I0,1X,ES24.16)', 'dirichlet ', rc, sum(coeffD)
  rc = pp_geom_destroy(geo)
  print '(A,I0)', 'destroy ', rc
end program layout_check
