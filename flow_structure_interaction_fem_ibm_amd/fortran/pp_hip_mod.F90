! pp_hip_mod.F90 -- ISO_C_BINDING interface of the MI355X pressure-Poisson
! library (include/pp_hip.h) for the Fortran flow solver.
!
! The derived types mirror the C structs field by field (BIND(C) layout);
! every routine returns the library's int32 status (PP_OK = 0).  Arrays are
! passed as C_LOC of the solver's module arrays in their own layout
! (0:nx+1, 0:ny+1, 0:nz+1), so no copies are made on the Fortran side.
!
! Replaces, as a drop-in behind solve_poisson's pp_solver_type switch
! (CORE_POISSON.F90:566), the MG branch CORE_POISSON.F90:638-676.
module pp_hip_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: PP_OK = 0, PP_NOT_CONVERGED = 1, PP_BAD_ARG = 2, &
       PP_BAD_GEOMETRY = 3, PP_HIP_ERROR = 4, PP_NONFINITE = 5, PP_NO_GEOMETRY = 6
  integer(c_int32_t), parameter, public :: PP_BC_NEUMANN = 2
  integer(c_int32_t), parameter, public :: PP_PREC_FP64 = 0, PP_PREC_MIXED = 1
  integer(c_int32_t), parameter, public :: PP_SMOOTH_LINE = 0, PP_SMOOTH_POINT_JACOBI = 1, &
                                            PP_SMOOTH_LINE_NATURAL = 2
  integer, parameter, public :: PP_K_CLASSES = 13
  integer(c_int32_t), parameter, public :: PP_BODY_ELLIPSOID = 3
  integer(c_int32_t), parameter, public :: PP_GEOM_ANALYTIC = 0, PP_GEOM_TRIANGULATED = 1
  ! advec_scheme (AMODULES.F90:60-62)
  integer(c_int32_t), parameter, public :: PP_ADAMS_BASHFORTH2 = 1, PP_CRANK_NICOLSON1 = 2, PP_CRANK_NICOLSON2 = 3

  type, bind(c), public :: pp_grid_desc
     integer(c_int32_t) :: nxc, nyc, nzc
     type(c_ptr) :: x, y, z            ! node coordinates x(1..nxc+1)
     integer(c_int32_t) :: bc(6)       ! pbcx1, pbcx2, pbcy1, pbcy2, pbcz1, pbcz2
     integer(c_int32_t) :: device
  end type pp_grid_desc

  type, bind(c), public :: pp_geometry
     type(c_ptr) :: iblank             ! INTEGER(1) (0:nx+1,0:ny+1,0:nz+1)
     type(c_ptr) :: fresh_cell         ! INTEGER(1) or C_NULL_PTR
     integer(c_int32_t) :: n_ghost
     type(c_ptr) :: i_ghost, j_ghost, k_ghost   ! INTEGER(4) (nGhost)
     type(c_ptr) :: i_cell, j_cell, k_cell      ! iCellIndex ... (nGhost)
     type(c_ptr) :: coeff              ! coeffGCMN(8, nGhost) REAL(8)
     type(c_ptr) :: probe_length       ! probeLength(nGhost) REAL(8)
  end type pp_geometry

  type, bind(c), public :: pp_solve_opts
     integer(c_int32_t) :: precision
     real(c_double) :: rel_tol
     real(c_double) :: abs_linf_tol    ! restol_Poisson
     integer(c_int32_t) :: max_outer   ! iterMax_Poisson
     integer(c_int32_t) :: cycles_per_outer
     integer(c_int32_t) :: nu_pre, nu_post
     real(c_double) :: omega
     integer(c_int32_t) :: mean_remove, project_rhs, use_graph, profile, krylov, nu_coarse, smoother
  end type pp_solve_opts

  ! one canonical body (canonical_body_type, xcent.., radiusx.., alpha, nPtsBodyMarker)
  type, bind(c), public :: pp_body_desc
     integer(c_int32_t) :: shape       ! PP_BODY_ELLIPSOID
     real(c_double) :: center(3), semi(3)
     real(c_double) :: alpha_deg
     integer(c_int32_t) :: n_theta
  end type pp_body_desc

  type, bind(c), public :: pp_solve_stats
     integer(c_int32_t) :: status, outer_iterations, vcycles
     real(c_double) :: rel_l2, abs_linf
     integer(c_int32_t) :: loc(3)
     real(c_double) :: b_l2, rhs_shift, mean_removed, ms, bytes
     real(c_double) :: history(64)
     real(c_double) :: kern_ms(PP_K_CLASSES), kern_bytes(PP_K_CLASSES)
     integer(c_int32_t) :: kern_launches(PP_K_CLASSES)
     real(c_double) :: kern_ms_l0(PP_K_CLASSES), kern_bytes_l0(PP_K_CLASSES)
     integer(c_int32_t) :: kern_launches_l0(PP_K_CLASSES)
  end type pp_solve_stats

  ! one UNSTRUCTURED_SURFACE body of unstruc_surface_in.dat: xBodyMarker.. as
  ! (3, nPtsBodyMarker), triElemNeig as (3, totNumTriElem), pointOutsideBody
  type, bind(c), public :: pp_surface_desc
     integer(c_int32_t) :: n_markers, n_elems
     type(c_ptr) :: xyz, tri
     real(c_double) :: outside(3)
  end type pp_surface_desc

  ! solve_ad: iterations, max(maxresu, maxresv, maxresw), the three, loc
  type, bind(c), public :: pp_ad_stats
     integer(c_int32_t) :: status, iterations
     real(c_double) :: max_residual
     real(c_double) :: residual_uvw(3)
     integer(c_int32_t) :: loc(3)
     real(c_double) :: ms
  end type pp_ad_stats

  public :: pp_default_opts, pp_last_error, pp_create, pp_destroy, pp_set_geometry, pp_solve, &
       pp_solve_device, pp_adjoint, pp_adjoint_device, pp_calc_residual, pp_ghost_rows, pp_mean_remove, pp_divergence, &
       pp_smooth, pp_null_vector, pp_level_info, pp_error_text, pp_geom_build, pp_geom_counts, pp_geom_get, &
       pp_geom_ghost_elements, pp_geom_destroy, pp_geom_dirichlet, pp_set_vel_ghosts, pp_correct_vel, &
       pp_correct_vel_device, pp_update_pressure, pp_update_pressure_device, pp_set_solve_ad, &
       pp_set_solve_ad_device, pp_solve_ad, pp_solve_ad_device, pp_calc_residual_ad, pp_geom_build_surfaces, &
       pp_geom_canonical_surface, pp_geom_fresh, pp_set_fresh_cells, pp_face_vel, pp_face_vel_device, &
       pp_solve_ad_cn2, pp_solve_ad_cn2_device, pp_geom_build_device, pp_geom_build_surfaces_device

  interface
     subroutine pp_default_opts(o) bind(c, name="pp_default_opts")
       import :: pp_solve_opts
       type(pp_solve_opts), intent(out) :: o
     end subroutine pp_default_opts

     function pp_last_error() bind(c, name="pp_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function pp_last_error

     function pp_create(ctx, g) bind(c, name="pp_create") result(rc)
       import :: c_ptr, c_int32_t, pp_grid_desc
       type(c_ptr), intent(out) :: ctx
       type(pp_grid_desc), intent(in) :: g
       integer(c_int32_t) :: rc
     end function pp_create

     function pp_destroy(ctx) bind(c, name="pp_destroy") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: rc
     end function pp_destroy

     function pp_set_geometry(ctx, geo) bind(c, name="pp_set_geometry") result(rc)
       import :: c_ptr, c_int32_t, pp_geometry
       type(c_ptr), value :: ctx
       type(pp_geometry), intent(in) :: geo
       integer(c_int32_t) :: rc
     end function pp_set_geometry

     function pp_solve(ctx, p_inout, rhs, bc_src, o, st) bind(c, name="pp_solve") result(rc)
       import :: c_ptr, c_int32_t, pp_solve_opts, pp_solve_stats
       type(c_ptr), value :: ctx, p_inout, rhs, bc_src
       type(pp_solve_opts), intent(in) :: o
       type(pp_solve_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_solve

     function pp_solve_device(ctx, d_p, d_rhs, d_bc, o, st) bind(c, name="pp_solve_device") result(rc)
       import :: c_ptr, c_int32_t, pp_solve_opts, pp_solve_stats
       type(c_ptr), value :: ctx, d_p, d_rhs, d_bc
       type(pp_solve_opts), intent(in) :: o
       type(pp_solve_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_solve_device

     function pp_adjoint(ctx, w, lambda, dj_dg, o, st) bind(c, name="pp_adjoint") result(rc)
       import :: c_ptr, c_int32_t, pp_solve_opts, pp_solve_stats
       type(c_ptr), value :: ctx, w, lambda, dj_dg
       type(pp_solve_opts), intent(in) :: o
       type(pp_solve_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_adjoint

     function pp_adjoint_device(ctx, w, lambda, dj_dg, o, st) bind(c, name="pp_adjoint_device") result(rc)
       import :: c_ptr, c_int32_t, pp_solve_opts, pp_solve_stats
       type(c_ptr), value :: ctx, w, lambda, dj_dg
       type(pp_solve_opts), intent(in) :: o
       type(pp_solve_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_adjoint_device

     function pp_calc_residual(ctx, var, r, res, resm, loc, l2) bind(c, name="pp_calc_residual") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, var, r, res
       real(c_double), intent(out) :: resm, l2
       integer(c_int32_t), intent(out) :: loc(3)
       integer(c_int32_t) :: rc
     end function pp_calc_residual

     function pp_ghost_rows(ctx, p, bc_src) bind(c, name="pp_ghost_rows") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx, p, bc_src
       integer(c_int32_t) :: rc
     end function pp_ghost_rows

     function pp_mean_remove(ctx, p, mean) bind(c, name="pp_mean_remove") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, p
       real(c_double), intent(out) :: mean
       integer(c_int32_t) :: rc
     end function pp_mean_remove

     function pp_divergence(ctx, fu, fv, fw, dtinv, nlu0, sum) bind(c, name="pp_divergence") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, fu, fv, fw, nlu0
       real(c_double), value :: dtinv
       real(c_double), intent(out) :: sum
       integer(c_int32_t) :: rc
     end function pp_divergence

     function pp_smooth(ctx, smoother, omega, sweeps, var, r) bind(c, name="pp_smooth") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, var, r
       integer(c_int32_t), value :: smoother, sweeps
       real(c_double), value :: omega
       integer(c_int32_t) :: rc
     end function pp_smooth

     function pp_null_vector(ctx, w) bind(c, name="pp_null_vector") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx, w
       integer(c_int32_t) :: rc
     end function pp_null_vector

     function pp_level_info(ctx, n_levels, dims) bind(c, name="pp_level_info") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: n_levels
       integer(c_int32_t), intent(out) :: dims(*)
       integer(c_int32_t) :: rc
     end function pp_level_info

     ! set_boundary replacement (host only): bodies -> iblank, fresh cells,
     ! ghost lists and GCM coefficients, read back with pp_geom_get
     function pp_geom_build(geo, g, n_body, bodies, mode, probe_norm, iblank_prev) &
          bind(c, name="pp_geom_build") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_grid_desc, pp_body_desc
       type(c_ptr), intent(out) :: geo
       type(pp_grid_desc), intent(in) :: g
       integer(c_int32_t), value :: n_body, mode
       type(pp_body_desc), intent(in) :: bodies(*)
       real(c_double), value :: probe_norm
       type(c_ptr), value :: iblank_prev
       integer(c_int32_t) :: rc
     end function pp_geom_build

     function pp_geom_counts(geo, n_ghost, n_blank, n_fresh) bind(c, name="pp_geom_counts") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: geo
       integer(c_int32_t), intent(out) :: n_ghost, n_blank, n_fresh
       integer(c_int32_t) :: rc
     end function pp_geom_counts

     function pp_geom_get(geo, iblank, fresh, ig, jg, kg, ic, jc, kc, coeff, probe, bi, ip, nrm) &
          bind(c, name="pp_geom_get") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: geo, iblank, fresh, ig, jg, kg, ic, jc, kc, coeff, probe, bi, ip, nrm
       integer(c_int32_t) :: rc
     end function pp_geom_get

     function pp_geom_ghost_elements(geo, body, elem) bind(c, name="pp_geom_ghost_elements") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: geo, body, elem
       integer(c_int32_t) :: rc
     end function pp_geom_ghost_elements

     function pp_geom_dirichlet(geo, coeffD) bind(c, name="pp_geom_dirichlet") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: geo, coeffD
       integer(c_int32_t) :: rc
     end function pp_geom_dirichlet

     ! fractional step either side of the solve: correct_vel and update_pressure
     function pp_set_vel_ghosts(ctx, coeffD, bi_weight, ip_weight) bind(c, name="pp_set_vel_ghosts") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, coeffD
       real(c_double), value :: bi_weight, ip_weight
       integer(c_int32_t) :: rc
     end function pp_set_vel_ghosts

     function pp_correct_vel(ctx, dt, p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw, vel_bi) &
          bind(c, name="pp_correct_vel") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw, vel_bi
       real(c_double), value :: dt
       integer(c_int32_t) :: rc
     end function pp_correct_vel

     function pp_correct_vel_device(ctx, dt, p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw, vel_bi) &
          bind(c, name="pp_correct_vel_device") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw, vel_bi
       real(c_double), value :: dt
       integer(c_int32_t) :: rc
     end function pp_correct_vel_device

     function pp_update_pressure(ctx, van_kan, pprime, p, bc_src) bind(c, name="pp_update_pressure") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx, pprime, p, bc_src
       integer(c_int32_t), value :: van_kan
       integer(c_int32_t) :: rc
     end function pp_update_pressure

     function pp_update_pressure_device(ctx, van_kan, pprime, p, bc_src) &
          bind(c, name="pp_update_pressure_device") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx, pprime, p, bc_src
       integer(c_int32_t), value :: van_kan
       integer(c_int32_t) :: rc
     end function pp_update_pressure_device

     ! set_solve_ad: viscTot and bcx/y/zvisc, reference layout
     function pp_set_solve_ad(ctx, dt, visc_tot, bcxvisc, bcyvisc, bczvisc) &
          bind(c, name="pp_set_solve_ad") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, visc_tot, bcxvisc, bcyvisc, bczvisc
       real(c_double), value :: dt
       integer(c_int32_t) :: rc
     end function pp_set_solve_ad

     function pp_set_solve_ad_device(ctx, dt, visc_tot, bcxvisc, bcyvisc, bczvisc) &
          bind(c, name="pp_set_solve_ad_device") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, visc_tot, bcxvisc, bcyvisc, bczvisc
       real(c_double), value :: dt
       integer(c_int32_t) :: rc
     end function pp_set_solve_ad_device

     ! solve_ad, GCM branch: u, v, w in/out; nlu, nlv, nlw; face_u/v/w; vel_bi (3, nGhost) or C_NULL_PTR
     function pp_solve_ad(ctx, advec_scheme, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, vel_bi, &
          itermax_ad, restol_ad, st) bind(c, name="pp_solve_ad") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_ad_stats
       type(c_ptr), value :: ctx, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, vel_bi
       integer(c_int32_t), value :: advec_scheme, itermax_ad
       real(c_double), value :: restol_ad
       type(pp_ad_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_solve_ad

     function pp_solve_ad_device(ctx, advec_scheme, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, vel_bi, &
          itermax_ad, restol_ad, st) bind(c, name="pp_solve_ad_device") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_ad_stats
       type(c_ptr), value :: ctx, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, vel_bi
       integer(c_int32_t), value :: advec_scheme, itermax_ad
       real(c_double), value :: restol_ad
       type(pp_ad_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_solve_ad_device

     ! solve_ad with CRANK_NICOLSON2: face_vel after every ghost update; face_u/v/w in/out;
     ! van_kan = 1 for frac_step_type VAN_KAN (p is then read), bcxu/bcyv/bczw the outer faces
     function pp_solve_ad_cn2(ctx, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, van_kan, p, bcxu, bcyv, &
          bczw, vel_bi, itermax_ad, restol_ad, st) bind(c, name="pp_solve_ad_cn2") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_ad_stats
       type(c_ptr), value :: ctx, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, p, bcxu, bcyv, bczw, vel_bi
       integer(c_int32_t), value :: van_kan, itermax_ad
       real(c_double), value :: restol_ad
       type(pp_ad_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_solve_ad_cn2

     function pp_solve_ad_cn2_device(ctx, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, van_kan, p, bcxu, &
          bcyv, bczw, vel_bi, itermax_ad, restol_ad, st) bind(c, name="pp_solve_ad_cn2_device") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_ad_stats
       type(c_ptr), value :: ctx, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, p, bcxu, bcyv, bczw, vel_bi
       integer(c_int32_t), value :: van_kan, itermax_ad
       real(c_double), value :: restol_ad
       type(pp_ad_stats), intent(out) :: st
       integer(c_int32_t) :: rc
     end function pp_solve_ad_cn2_device

     ! face_vel: face_u/v/w from u, v, w (uTilde with p for VAN_KAN), reference layout
     function pp_face_vel(ctx, van_kan, dt, p, u, v, w, bcxu, bcyv, bczw, face_u, face_v, face_w) &
          bind(c, name="pp_face_vel") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, p, u, v, w, bcxu, bcyv, bczw, face_u, face_v, face_w
       integer(c_int32_t), value :: van_kan
       real(c_double), value :: dt
       integer(c_int32_t) :: rc
     end function pp_face_vel

     function pp_face_vel_device(ctx, van_kan, dt, p, u, v, w, bcxu, bcyv, bczw, face_u, face_v, face_w) &
          bind(c, name="pp_face_vel_device") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, p, u, v, w, bcxu, bcyv, bczw, face_u, face_v, face_w
       integer(c_int32_t), value :: van_kan
       real(c_double), value :: dt
       integer(c_int32_t) :: rc
     end function pp_face_vel_device

     ! fresh cells of the momentum solve (GCM_set_freshcell's tables, as pp_geom_fresh gives them)
     function pp_set_fresh_cells(ctx, n_fresh, i_fresh, j_fresh, k_fresh, i_cell, j_cell, k_cell, coeffD) &
          bind(c, name="pp_set_fresh_cells") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx, i_fresh, j_fresh, k_fresh, i_cell, j_cell, k_cell, coeffD
       integer(c_int32_t), value :: n_fresh
       integer(c_int32_t) :: rc
     end function pp_set_fresh_cells

     ! calc_residual_ad of one field: resm, loc (1-based), res (or C_NULL_PTR)
     function pp_calc_residual_ad(ctx, advec_scheme, var, r, face_u, face_v, face_w, res, resm, loc) &
          bind(c, name="pp_calc_residual_ad") result(rc)
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, var, r, face_u, face_v, face_w, res
       integer(c_int32_t), value :: advec_scheme
       real(c_double), intent(out) :: resm
       integer(c_int32_t), intent(out) :: loc(3)
       integer(c_int32_t) :: rc
     end function pp_calc_residual_ad

     function pp_geom_build_surfaces(geo, g, n_surface, surfaces, probe_length_normalized, iblank_prev) &
          bind(c, name="pp_geom_build_surfaces") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_grid_desc, pp_surface_desc
       type(c_ptr), intent(out) :: geo
       type(pp_grid_desc), intent(in) :: g
       integer(c_int32_t), value :: n_surface
       type(pp_surface_desc), intent(in) :: surfaces(*)
       real(c_double), value :: probe_length_normalized
       type(c_ptr), value :: iblank_prev
       integer(c_int32_t) :: rc
     end function pp_geom_build_surfaces

     ! pp_geom_build (triangulated) and pp_geom_build_surfaces with the per-cell and per-ghost work on GPU
     ! g%device; the same handle and getters, the same results bit for bit
     function pp_geom_build_device(geo, g, n_body, bodies, probe_norm, iblank_prev) &
          bind(c, name="pp_geom_build_device") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_grid_desc, pp_body_desc
       type(c_ptr), intent(out) :: geo
       type(pp_grid_desc), intent(in) :: g
       integer(c_int32_t), value :: n_body
       type(pp_body_desc), intent(in) :: bodies(*)
       real(c_double), value :: probe_norm
       type(c_ptr), value :: iblank_prev
       integer(c_int32_t) :: rc
     end function pp_geom_build_device

     function pp_geom_build_surfaces_device(geo, g, n_surface, surfaces, probe_length_normalized, iblank_prev) &
          bind(c, name="pp_geom_build_surfaces_device") result(rc)
       import :: c_ptr, c_int32_t, c_double, pp_grid_desc, pp_surface_desc
       type(c_ptr), intent(out) :: geo
       type(pp_grid_desc), intent(in) :: g
       integer(c_int32_t), value :: n_surface
       type(pp_surface_desc), intent(in) :: surfaces(*)
       real(c_double), value :: probe_length_normalized
       type(c_ptr), value :: iblank_prev
       integer(c_int32_t) :: rc
     end function pp_geom_build_surfaces_device

     ! counts with xyz = tri = C_NULL_PTR, then the arrays
     function pp_geom_canonical_surface(body, n_markers, n_elems, xyz, tri) &
          bind(c, name="pp_geom_canonical_surface") result(rc)
       import :: c_ptr, c_int32_t, pp_body_desc
       type(pp_body_desc), intent(in) :: body
       integer(c_int32_t), intent(inout) :: n_markers, n_elems
       type(c_ptr), value :: xyz, tri
       integer(c_int32_t) :: rc
     end function pp_geom_canonical_surface

     ! GCM_set_freshcell for the fresh cells of a geometry built with iblank_prev (n_fresh of
     ! pp_geom_counts): cells, stencil corners, coeffGCMFreshD(8, n), intercepts (3, n), body, element
     function pp_geom_fresh(geo, i_fresh, j_fresh, k_fresh, i_cell, j_cell, k_cell, coeffD, bi, body, elem) &
          bind(c, name="pp_geom_fresh") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: geo, i_fresh, j_fresh, k_fresh, i_cell, j_cell, k_cell, coeffD, bi, body, elem
       integer(c_int32_t) :: rc
     end function pp_geom_fresh

     function pp_geom_destroy(geo) bind(c, name="pp_geom_destroy") result(rc)
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: geo
       integer(c_int32_t) :: rc
     end function pp_geom_destroy
  end interface

contains

  ! pp_last_error() as a Fortran string (for the solver's PRINT statements)
  function pp_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
       function c_strlen(str) bind(c, name="strlen") result(n)
         import :: c_ptr, c_size_t
         type(c_ptr), value :: str
         integer(c_size_t) :: n
       end function c_strlen
    end interface
    p = pp_last_error()
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function pp_error_text

end module pp_hip_mod
