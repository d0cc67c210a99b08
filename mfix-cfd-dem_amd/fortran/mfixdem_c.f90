!----------------------------------------------------------------------!
! mfixdem_c: ISO_C_BINDING interface to the gfx950 DEM library        !
! (include/mfixdem.h).  MFiX-side UDF overrides of the DEM routines    !
! USE this module and call these entry points instead of running the   !
! CPU loops.  Types and names mirror the C header.                     !
!----------------------------------------------------------------------!
module mfixdem_c
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MFIXDEM_ABI_VERSION = 3
  integer, parameter :: MFIXDEM_MAX_PHASES = 4
  integer, parameter :: MFIXDEM_MAX_PAIRS_IN = 10
  real(c_double), parameter :: MFIXDEM_UNDEFINED = 9.87654321d31

  integer(c_int32_t), parameter :: MFIXDEM_OK = 0
  integer(c_int32_t), parameter :: MFIXDEM_LSD = 0, MFIXDEM_HERTZ = 1
  integer(c_int32_t), parameter :: MFIXDEM_EULER = 0, MFIXDEM_ADAMS_BASHFORTH = 1
  integer(c_int32_t), parameter :: MFIXDEM_FP64 = 0, MFIXDEM_MIXED = 1
  integer(c_int32_t), parameter :: MFIXDEM_DRAG_GIDASPOW = 0, MFIXDEM_DRAG_WEN_YU = 1
  integer(c_int32_t), parameter :: MFIXDEM_INTERP_GARG_2012 = 0, MFIXDEM_INTERP_NONE = 1
  integer(c_int32_t), parameter :: MFIXDEM_INTERP_SQUARE_DPVM = 2, MFIXDEM_INTERP_GAUSS_DPVM = 3
  integer(c_int32_t), parameter :: MFIXDEM_INTERP_LINEAR_HAT = 4
  integer(c_int32_t), parameter :: MFIXDEM_ROWS_THREAD = 0, MFIXDEM_ROWS_WAVE128 = 1
  integer(c_int32_t), parameter :: MFIXDEM_ROWS_WAVE192 = 2, MFIXDEM_ROWS_WAVE256 = 3
  integer(c_int32_t), parameter :: MFIXDEM_NSEARCH_NSQUARE = 1, MFIXDEM_NSEARCH_GRID = 4
  integer(c_int32_t), parameter :: MFIXDEM_DEBUG_STORE_FORCES = 1
  integer(c_int32_t), parameter :: MFIXDEM_BC_MASS_INFLOW = 1, MFIXDEM_BC_MASS_OUTFLOW = 2
  integer(c_int32_t), parameter :: MFIXDEM_BC_OUTFLOW = 3
  integer(c_int32_t), parameter :: MFIXDEM_PLANE_W = 0, MFIXDEM_PLANE_E = 1
  integer(c_int32_t), parameter :: MFIXDEM_PLANE_S = 2, MFIXDEM_PLANE_N = 3
  integer(c_int32_t), parameter :: MFIXDEM_PLANE_B = 4, MFIXDEM_PLANE_T = 5

  type, bind(C) :: mfixdem_params
     integer(c_int32_t) :: abi_version, coll_model, integrator, precision
     integer(c_int32_t) :: nphases, neighbor_search_n, coupled, drag_type
     integer(c_int32_t) :: interp_scheme, model_b, max_pip, imax, jmax, kmax
     integer(c_int32_t) :: desgridsearch_imax, desgridsearch_jmax, desgridsearch_kmax
     integer(c_int32_t) :: remove_rogue, des_interp_on, des_interp_mean_fields
     real(c_double) :: x_min, y_min, z_min, xlength, ylength, zlength
     real(c_double) :: dtsolid, dtsolid_factor
     real(c_double) :: grav(3)
     real(c_double) :: factor_rlm, kn, kn_w, kt_fac, kt_w_fac
     real(c_double) :: des_etat_fac, des_etat_w_fac, mew, mew_w
     real(c_double) :: des_en_input(MFIXDEM_MAX_PAIRS_IN), des_et_input(MFIXDEM_MAX_PAIRS_IN)
     real(c_double) :: des_en_wall_input(MFIXDEM_MAX_PHASES), des_et_wall_input(MFIXDEM_MAX_PHASES)
     real(c_double) :: e_young(MFIXDEM_MAX_PHASES), v_poisson(MFIXDEM_MAX_PHASES)
     real(c_double) :: ew_young, vw_poisson
     real(c_double) :: d_p0(MFIXDEM_MAX_PHASES), ro_s0(MFIXDEM_MAX_PHASES)
     real(c_double) :: max_radius, ro_g0, mu_g0, des_epg_clip
     real(c_double) :: des_interp_width
     integer(c_int32_t) :: des_neighbor_search, des_explicitly_coupled
     type(c_ptr) :: dx, dy, dz      ! c_loc(DX(IMIN1)) ... or c_null_ptr (uniform)
  end type mfixdem_params

  ! one DEM boundary of mfix.dat: BC_TYPE, BC_PLANE, BC_X_w ... BC_Z_t,
  ! BC_U_s/V_s/W_s, BC_EP_s, BC_MASSFLOW_s, BC_MI_START_TIME/END_TIME
  type, bind(C) :: mfixdem_dem_bc
     integer(c_int32_t) :: type, plane
     real(c_double) :: x_w, x_e, y_s, y_n, z_b, z_t
     real(c_double) :: u_s(MFIXDEM_MAX_PHASES), v_s(MFIXDEM_MAX_PHASES), w_s(MFIXDEM_MAX_PHASES)
     real(c_double) :: ep_s(MFIXDEM_MAX_PHASES), massflow_s(MFIXDEM_MAX_PHASES)
     real(c_double) :: mi_start_time, mi_end_time
  end type mfixdem_dem_bc

  type, bind(C) :: mfixdem_status
     integer(c_int64_t) :: steps, rebuilds, n_particles, n_pairs, n_contacts
     integer(c_int64_t) :: rogue_deleted, epg_errors, zero_distance, out_of_domain
  end type mfixdem_status

  interface
     integer(c_int32_t) function mfixdem_params_default(p) bind(C, name="mfixdem_params_default")
       import :: c_int32_t, mfixdem_params
       type(mfixdem_params), intent(out) :: p
     end function
     integer(c_int32_t) function mfixdem_create(p, h) bind(C, name="mfixdem_create")
       import :: c_int32_t, c_ptr, mfixdem_params
       type(mfixdem_params), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function
     integer(c_int32_t) function mfixdem_destroy(h) bind(C, name="mfixdem_destroy")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_set_particles(h, n, pos, vel, omega, radius, ro_sol, &
          phase, ld) bind(C, name="mfixdem_set_particles")
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: n, ld
       real(c_double), intent(in) :: pos(*), vel(*), omega(*), radius(*), ro_sol(*)
       integer(c_int32_t), intent(in) :: phase(*)
     end function
     integer(c_int32_t) function mfixdem_get_particles(h, pos, vel, omega, state, ld) &
          bind(C, name="mfixdem_get_particles")
       import :: c_int32_t, c_int64_t, c_ptr, c_double, c_int8_t
       type(c_ptr), value :: h
       real(c_double), intent(out) :: pos(*), vel(*), omega(*)
       integer(c_int8_t), intent(out) :: state(*)
       integer(c_int64_t), value :: ld
     end function
     integer(c_int32_t) function mfixdem_set_gas(h, u_g, v_g, w_g, p_force) &
          bind(C, name="mfixdem_set_gas")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u_g(*), v_g(*), w_g(*), p_force(*)
     end function
     integer(c_int32_t) function mfixdem_get_gas(h, ep_g, rop_s) bind(C, name="mfixdem_get_gas")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: ep_g(*), rop_s(*)
     end function
     integer(c_int32_t) function mfixdem_init(h) bind(C, name="mfixdem_init")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_step(h, nn) bind(C, name="mfixdem_step")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nn
     end function
     integer(c_int32_t) function mfixdem_run(h, nn_first, nsteps) bind(C, name="mfixdem_run")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nn_first, nsteps
     end function
     integer(c_int32_t) function mfixdem_synchronize(h) bind(C, name="mfixdem_synchronize")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_calc_force_dem(h) bind(C, name="mfixdem_calc_force_dem")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_calc_drag_des(h) bind(C, name="mfixdem_calc_drag_des")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_cfnewvalues(h) bind(C, name="mfixdem_cfnewvalues")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_desgrid_pic(h) bind(C, name="mfixdem_desgrid_pic")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_neighbour(h) bind(C, name="mfixdem_neighbour")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_particles_in_cell(h) bind(C, name="mfixdem_particles_in_cell")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_set_gas_pressure(h, p_g) &
          bind(C, name="mfixdem_set_gas_pressure")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: p_g(*)
     end function
     integer(c_int32_t) function mfixdem_get_p_force(h, p_force) &
          bind(C, name="mfixdem_get_p_force")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: p_force(*)
     end function
     integer(c_int32_t) function mfixdem_calc_interp_weights(h) &
          bind(C, name="mfixdem_calc_interp_weights")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_comp_mean_fields(h) bind(C, name="mfixdem_comp_mean_fields")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     ! DRAG_GS_GAS0 (drag_gs_des0.f:235): F_GDS, DRAG_AM, DRAG_BM for the gas solver
     integer(c_int32_t) function mfixdem_drag_gs_gas(h) bind(C, name="mfixdem_drag_gs_gas")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     ! PARTICLE_STATE holes (0) before init
     integer(c_int32_t) function mfixdem_set_particle_state(h, state, n) &
          bind(C, name="mfixdem_set_particle_state")
       import :: c_int32_t, c_int64_t, c_int8_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int8_t), intent(in) :: state(*)
       integer(c_int64_t), value :: n
     end function
     ! STL walls: VERTEX(3,3,nf) and NORM_FACE(3,nf) as GET_STL_DATA reads them
     integer(c_int32_t) function mfixdem_set_facets(h, nf, vertex, normal, default_walls) &
          bind(C, name="mfixdem_set_facets")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nf
       real(c_double), intent(in) :: vertex(3, 3, *), normal(3, *)
       integer(c_int32_t), value :: default_walls
     end function
     ! BC_STL wall motion: BC_WALL_VEL, BC_WALL_OMEGA (rad/s), BC_WALL_ROT_CENTER
     integer(c_int32_t) function mfixdem_set_wall_motion(h, first, count, vel, omega, &
          rot_center) bind(C, name="mfixdem_set_wall_motion")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: first, count
       real(c_double), intent(in) :: vel(3), omega(3), rot_center(3)
     end function
     ! DEM mass inlets and outlets, before mfixdem_init
     integer(c_int32_t) function mfixdem_set_dem_bcs(h, n, bcs, seed) &
          bind(C, name="mfixdem_set_dem_bcs")
       import :: c_int32_t, c_int64_t, c_ptr, mfixdem_dem_bc
       type(c_ptr), value :: h
       integer(c_int32_t), value :: n
       type(mfixdem_dem_bc), intent(in) :: bcs(*)
       integer(c_int64_t), value :: seed
     end function
     ! DES_TIME_INIT's coupled part with DES_EXPLICITLY_COUPLED (DRAG_FC)
     integer(c_int32_t) function mfixdem_des_time_init(h) bind(C, name="mfixdem_des_time_init")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int32_t) function mfixdem_get_gas_drag(h, f_gds, drag_am, drag_bm) &
          bind(C, name="mfixdem_get_gas_drag")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: f_gds(*), drag_am(*), drag_bm(*)
     end function
     integer(c_int32_t) function mfixdem_set_debug(h, flags) bind(C, name="mfixdem_set_debug")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: flags
     end function
     integer(c_int32_t) function mfixdem_get_forces(h, fc, tow, ld) bind(C, name="mfixdem_get_forces")
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: fc(*), tow(*)
       integer(c_int64_t), value :: ld
     end function
     integer(c_int32_t) function mfixdem_get_pairs(h, neighbor_index, neighbors, pft, capacity, &
          npairs) bind(C, name="mfixdem_get_pairs")
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: neighbor_index(*), neighbors(*)
       real(c_double), intent(out) :: pft(*)
       integer(c_int64_t), value :: capacity
       integer(c_int64_t), intent(out) :: npairs
     end function
     ! restart of NEIGHBORS, NEIGHBOR_INDEX, PFT_NEIGHBOR (read_res1_des_mod.f:108)
     integer(c_int32_t) function mfixdem_set_pairs(h, neighbor_index, neighbors, pft, npairs) &
          bind(C, name="mfixdem_set_pairs")
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), intent(in) :: neighbor_index(*), neighbors(*)
       real(c_double), intent(in) :: pft(*)
       integer(c_int64_t), value :: npairs
     end function
     integer(c_int32_t) function mfixdem_get_cells(h, dg_pijk, pijk, ld) &
          bind(C, name="mfixdem_get_cells")
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: dg_pijk(*), pijk(*)
       integer(c_int64_t), value :: ld
     end function
     integer(c_int32_t) function mfixdem_set_tuning(h, cell_scale, row_kernel) &
          bind(C, name="mfixdem_set_tuning")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: cell_scale
       integer(c_int32_t), value :: row_kernel
     end function
     integer(c_int32_t) function mfixdem_get_device_bytes(h, bytes) &
          bind(C, name="mfixdem_get_device_bytes")
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int64_t), intent(out) :: bytes
     end function
     integer(c_int32_t) function mfixdem_get_status(h, st) bind(C, name="mfixdem_get_status")
       import :: c_int32_t, c_ptr, mfixdem_status
       type(c_ptr), value :: h
       type(mfixdem_status), intent(out) :: st
     end function
  end interface

contains

  ! Map a nonzero status to the reference's fatal path (LOG_ERROR in MFiX).
  subroutine mfixdem_check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= MFIXDEM_OK) then
       write(*,'(a,a,a,i0)') 'mfixdem: ', where, ' failed with status ', rc
       error stop 1
    endif
  end subroutine mfixdem_check

end module mfixdem_c
