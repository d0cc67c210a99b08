!----------------------------------------------------------------------!
! mfixdem_bridge: the libmfixdem handle of an MFiX case.  A case adds   !
! this file and the routine overrides beside it (same file and module  !
! names as the model files they replace) to its directory; MFiX's build !
! compiles them in place of model/des/*.f.  mfixdem_open maps the DEM   !
! namelist state of the MFiX modules onto mfixdem_params and hands the  !
! particles (and the gas field when coupled) to the GPU; the overrides  !
! then run each DES_TIME_STEP routine on the device-resident state, and !
! mfixdem_pull brings the particle arrays back when MFiX code outside   !
! the DEM loop needs them (output, restart, user hooks).               !
!                                                                      !
! Phases: MFiX numbers the DEM phases MMAX+1 .. MMAX+DES_MMAX after the !
! continuum ones; the library numbers them 1 .. DES_MMAX.              !
!----------------------------------------------------------------------!
module mfixdem_bridge
  use, intrinsic :: iso_c_binding
  use mfixdem_c
  implicit none
  private
  public :: h, mfixdem_open, mfixdem_close, mfixdem_pull, mfixdem_push_gas, &
       mfixdem_push_pressure, mfixdem_fail

  type(c_ptr), save :: h = c_null_ptr
  ! cell sizes of the interior cells, kept for mfixdem_create
  real(c_double), allocatable, target, save :: cell_dx(:), cell_dy(:), cell_dz(:)

contains

  ! The library's error text on MFiX's fatal path (an MFiX build replaces
  ! the error stop with LOG_ERROR(), error_manager_mod.f:249).
  subroutine mfixdem_fail(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc == MFIXDEM_OK) return
    write(*, '(a,a,a,i0)') 'libmfixdem: ', what, ' failed with status ', rc
    error stop 1
  end subroutine mfixdem_fail

  ! Once, after MAKE_ARRAYS_DES (make_arrays_des.f) has set up the particles.
  subroutine mfixdem_open()
    use param1, only: UNDEFINED, UNDEFINED_I
    use constant, only: GRAVITY_X, GRAVITY_Y, GRAVITY_Z
    use geometry, only: IMAX, JMAX, KMAX, XLENGTH, YLENGTH, ZLENGTH, X_MIN, Y_MIN, Z_MIN, &
         DX, DY, DZ, IMIN1, JMIN1, KMIN1
    use physprop, only: MMAX, D_P0, RO_S0, RO_G0, MU_G0
    use run, only: DES_CONTINUUM_COUPLED, MODEL_B
    use derived_types, only: DRAG_TYPE_ENUM, GIDASPOW, WEN_YU
    use particle_filter, only: DES_INTERP_SCHEME_ENUM, DES_INTERP_NONE, DES_INTERP_GARG, &
         DES_INTERP_DPVM, DES_INTERP_GAUSS, DES_INTERP_LHAT, DES_INTERP_ON, &
         DES_INTERP_MEAN_FIELDS, DES_INTERP_WIDTH
    use discretelement, only: PIP, MAX_PIP, DES_MMAX, DES_COLL_MODEL_ENUM, HERTZIAN, &
         DES_INTG_METHOD, DTSOLID, NEIGHBOR_SEARCH_N, DES_NEIGHBOR_SEARCH, FACTOR_RLM, &
         DESGRIDSEARCH_IMAX, DESGRIDSEARCH_JMAX, DESGRIDSEARCH_KMAX, REMOVE_ROGUE_PARTICLES, &
         KN, KN_W, KT_FAC, KT_W_FAC, DES_ETAT_FAC, DES_ETAT_W_FAC, MEW, MEW_W, &
         DES_EN_INPUT, DES_ET_INPUT, DES_EN_WALL_INPUT, DES_ET_WALL_INPUT, &
         E_YOUNG, EW_YOUNG, V_POISSON, VW_POISSON, MAX_RADIUS, DES_EPG_CLIP, &
         DES_EXPLICITLY_COUPLED, DES_POS_NEW, DES_VEL_NEW, OMEGA_NEW, DES_RADIUS, RO_SOL, &
         PIJK, PARTICLE_STATE
    type(mfixdem_params) :: p
    integer :: m, npair
    integer(c_int32_t), allocatable :: phase(:)

    call mfixdem_fail(mfixdem_params_default(p), 'params_default')
    if (DES_MMAX > MFIXDEM_MAX_PHASES) call mfixdem_fail(9_c_int32_t, 'DES_MMAX > 4')
    p%nphases = DES_MMAX
    p%max_pip = MAX_PIP
    p%coll_model = merge(MFIXDEM_HERTZ, MFIXDEM_LSD, DES_COLL_MODEL_ENUM == HERTZIAN)
    p%integrator = merge(MFIXDEM_ADAMS_BASHFORTH, MFIXDEM_EULER, &
         trim(DES_INTG_METHOD) == 'ADAMS_BASHFORTH')
    p%dtsolid = DTSOLID
    p%grav = [GRAVITY_X, GRAVITY_Y, GRAVITY_Z]
    p%neighbor_search_n = NEIGHBOR_SEARCH_N
    p%des_neighbor_search = DES_NEIGHBOR_SEARCH
    p%factor_rlm = FACTOR_RLM
    p%desgridsearch_imax = merge(0, DESGRIDSEARCH_IMAX, DESGRIDSEARCH_IMAX == UNDEFINED_I)
    p%desgridsearch_jmax = merge(0, DESGRIDSEARCH_JMAX, DESGRIDSEARCH_JMAX == UNDEFINED_I)
    p%desgridsearch_kmax = merge(0, DESGRIDSEARCH_KMAX, DESGRIDSEARCH_KMAX == UNDEFINED_I)
    p%remove_rogue = merge(1, 0, REMOVE_ROGUE_PARTICLES)
    ! collision parameters (check_solids_dem.f): LSD stiffnesses, the
    ! tangential factors, friction, restitution tables and Hertz moduli
    p%kn = KN; p%kn_w = KN_W
    p%kt_fac = KT_FAC; p%kt_w_fac = KT_W_FAC
    p%des_etat_fac = DES_ETAT_FAC; p%des_etat_w_fac = DES_ETAT_W_FAC
    p%mew = MEW; p%mew_w = MEW_W
    npair = DES_MMAX * (DES_MMAX + 1) / 2       ! upper triangle, (1,1), (1,2), ...
    p%des_en_input(1:npair) = DES_EN_INPUT(1:npair)
    p%des_et_input(1:npair) = DES_ET_INPUT(1:npair)
    p%ew_young = EW_YOUNG; p%vw_poisson = VW_POISSON
    do m = 1, DES_MMAX
       p%des_en_wall_input(m) = DES_EN_WALL_INPUT(m)
       p%des_et_wall_input(m) = DES_ET_WALL_INPUT(m)
       p%e_young(m) = E_YOUNG(MMAX + m)
       p%v_poisson(m) = V_POISSON(MMAX + m)
       p%d_p0(m) = D_P0(MMAX + m)
       p%ro_s0(m) = RO_S0(MMAX + m)
    enddo
    p%max_radius = MAX_RADIUS
    ! fluid grid: the interior cells DX(IMIN1 .. IMIN1+IMAX-1) after SHIFT_DXYZ
    p%imax = IMAX; p%jmax = JMAX; p%kmax = KMAX
    p%x_min = X_MIN; p%y_min = Y_MIN; p%z_min = Z_MIN
    p%xlength = XLENGTH; p%ylength = YLENGTH; p%zlength = ZLENGTH
    cell_dx = DX(IMIN1:IMIN1 + IMAX - 1)
    cell_dy = DY(JMIN1:JMIN1 + JMAX - 1)
    cell_dz = DZ(KMIN1:KMIN1 + KMAX - 1)
    p%dx = c_loc(cell_dx); p%dy = c_loc(cell_dy); p%dz = c_loc(cell_dz)
    ! gas coupling
    p%coupled = merge(1, 0, DES_CONTINUUM_COUPLED)
    p%model_b = merge(1, 0, MODEL_B)
    p%des_explicitly_coupled = merge(1, 0, DES_EXPLICITLY_COUPLED)
    p%ro_g0 = RO_G0; p%mu_g0 = MU_G0
    p%des_epg_clip = DES_EPG_CLIP
    if (DES_CONTINUUM_COUPLED) then
       if (DRAG_TYPE_ENUM == GIDASPOW) then
          p%drag_type = MFIXDEM_DRAG_GIDASPOW
       else if (DRAG_TYPE_ENUM == WEN_YU) then
          p%drag_type = MFIXDEM_DRAG_WEN_YU
       else
          call mfixdem_fail(9_c_int32_t, 'DRAG_TYPE (GIDASPOW and WEN_YU only)')
       endif
       select case (DES_INTERP_SCHEME_ENUM)
       case (DES_INTERP_GARG);  p%interp_scheme = MFIXDEM_INTERP_GARG_2012
       case (DES_INTERP_NONE);  p%interp_scheme = MFIXDEM_INTERP_NONE
       case (DES_INTERP_DPVM);  p%interp_scheme = MFIXDEM_INTERP_SQUARE_DPVM
       case (DES_INTERP_GAUSS); p%interp_scheme = MFIXDEM_INTERP_GAUSS_DPVM
       case (DES_INTERP_LHAT);  p%interp_scheme = MFIXDEM_INTERP_LINEAR_HAT
       case default
          call mfixdem_fail(9_c_int32_t, 'DES_INTERP_SCHEME')
       end select
       p%des_interp_on = merge(1, 0, DES_INTERP_ON)
       p%des_interp_mean_fields = merge(1, 0, DES_INTERP_MEAN_FIELDS)
       p%des_interp_width = DES_INTERP_WIDTH
    endif
    call mfixdem_fail(mfixdem_create(p, h), 'create')

    ! particles: PIP slots of the (MAX_PIP, 3) arrays, holes included
    allocate(phase(max(PIP, 1)))
    phase = 1
    if (PIP > 0) phase(1:PIP) = PIJK(1:PIP, 5) - MMAX
    call mfixdem_fail(mfixdem_set_particles(h, int(PIP, c_int64_t), DES_POS_NEW, &
         DES_VEL_NEW, OMEGA_NEW, DES_RADIUS, RO_SOL, phase, int(MAX_PIP, c_int64_t)), &
         'set_particles')
    if (PIP > 0) then
       if (any(PARTICLE_STATE(1:PIP) == 0)) call mfixdem_fail(mfixdem_set_particle_state(h, &
            PARTICLE_STATE, int(PIP, c_int64_t)), 'set_particle_state')
    endif
    if (DES_CONTINUUM_COUPLED) call mfixdem_push_gas()
    call mfixdem_fail(mfixdem_init(h), 'init')
  end subroutine mfixdem_open

  ! The gas field of the current fluid step (after the gas solve).
  subroutine mfixdem_push_gas()
    use fldvar, only: U_G, V_G, W_G
    use discretelement, only: P_FORCE
    call mfixdem_fail(mfixdem_set_gas(h, U_G, V_G, W_G, P_FORCE), 'set_gas')
  end subroutine mfixdem_push_gas

  ! CALC_PG_GRAD on the device from the gas pressure; P_FORCE comes back
  ! for the fluid solver.
  subroutine mfixdem_push_pressure()
    use fldvar, only: P_G
    use discretelement, only: P_FORCE
    call mfixdem_fail(mfixdem_set_gas_pressure(h, P_G), 'set_gas_pressure')
    call mfixdem_fail(mfixdem_get_p_force(h, P_FORCE), 'get_p_force')
  end subroutine mfixdem_push_pressure

  ! Particle state, cells and mean fields back into the MFiX arrays.
  subroutine mfixdem_pull()
    use physprop, only: MMAX
    use fldvar, only: EP_G, ROP_S
    use run, only: DES_CONTINUUM_COUPLED
    use discretelement, only: MAX_PIP, DES_POS_NEW, DES_VEL_NEW, OMEGA_NEW, &
         PARTICLE_STATE, PIJK
    integer(c_int32_t), allocatable :: dg(:), pijk4(:, :)   ! PIJK(:,1:4) planes
    call mfixdem_fail(mfixdem_get_particles(h, DES_POS_NEW, DES_VEL_NEW, OMEGA_NEW, &
         PARTICLE_STATE, int(MAX_PIP, c_int64_t)), 'get_particles')
    allocate(dg(MAX_PIP), pijk4(MAX_PIP, 4))
    call mfixdem_fail(mfixdem_get_cells(h, dg, pijk4, int(MAX_PIP, c_int64_t)), 'get_cells')
    PIJK(:, 1:4) = pijk4
    ! ROP_S of the DEM phases: columns MMAX+1 .. MMAX+DES_MMAX, in sequence
    if (DES_CONTINUUM_COUPLED) &
         call mfixdem_fail(mfixdem_get_gas(h, EP_G, ROP_S(1, MMAX + 1)), 'get_gas')
  end subroutine mfixdem_pull

  subroutine mfixdem_close()
    if (c_associated(h)) call mfixdem_fail(mfixdem_destroy(h), 'destroy')
    h = c_null_ptr
  end subroutine mfixdem_close

end module mfixdem_bridge
