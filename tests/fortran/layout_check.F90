! Links the Fortran binding against the library and prints what the C side
! wrote into the BIND(C) types (no GPU needed): pp_default_opts, the status
! and error text of pp_create with a bad grid, and the type sizes.
program layout_check
  use iso_c_binding
  use pp_hip_mod
  implicit none
  type(pp_solve_opts) :: o
  type(pp_grid_desc) :: g
  type(pp_solve_stats) :: st
  type(c_ptr) :: ctx
  integer(c_int32_t) :: rc
  real(c_double), target :: xs(3)
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
end program layout_check
