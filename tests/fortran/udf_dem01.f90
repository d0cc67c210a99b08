!----------------------------------------------------------------------!
! udf_dem01: DEM01 (one sphere, d = 0.2 m, rho = 2600, dropped from    !
! y = 0.5 onto the floor of a 1 m box, LSD) run the way an MFiX build   !
! with the UDF overrides runs it: the MFiX module variables are set as  !
! the namelist and MAKE_ARRAYS_DES would leave them, the handle opens   !
! in INIT_PARTICLES_IN_CELL, and each DES_TIME_STEP calls the          !
! overridden routines in des_time_march.f's order.                     !
! Usage: udf_dem01 KN EN NSTEPS; prints "step y vy" per step.          !
!----------------------------------------------------------------------!
program udf_dem01
  use, intrinsic :: iso_c_binding
  use param1, only: UNDEFINED, UNDEFINED_I
  use constant
  use geometry
  use physprop
  use run
  use derived_types
  use particle_filter
  use discretelement
  use calc_force_dem_mod, only: CALC_FORCE_DEM, CALC_COORDINATION
  use calc_drag_des_mod, only: CALC_DRAG_DES
  use cfnewvalues_mod, only: CFNEWVALUES
  use neighbour_mod, only: NEIGHBOUR
  use particles_in_cell_mod, only: INIT_PARTICLES_IN_CELL
  use mfixdem_c, only: mfixdem_desgrid_pic, mfixdem_get_particles
  use mfixdem_bridge, only: h, mfixdem_fail, mfixdem_close
  implicit none
  integer :: nn, nsteps
  double precision :: kn_in, en_in
  character(len=64) :: arg

  kn_in = 1.0d4; en_in = 0.9d0; nsteps = 300
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg); read(arg, *) kn_in
  endif
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg); read(arg, *) en_in
  endif
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg); read(arg, *) nsteps
  endif

  ! geometry: 1 m box, one fluid cell, DX shifted to DX(IMIN1 = 2)
  IMAX = 1; JMAX = 1; KMAX = 1; IMIN1 = 2; JMIN1 = 2; KMIN1 = 2
  XLENGTH = 1.0d0; YLENGTH = 1.0d0; ZLENGTH = 1.0d0
  X_MIN = 0.0d0; Y_MIN = 0.0d0; Z_MIN = 0.0d0
  DX = 1.0d0; DY = 1.0d0; DZ = 1.0d0
  GRAVITY_X = 0.0d0; GRAVITY_Y = -9.80665d0; GRAVITY_Z = 0.0d0
  MMAX = 0; D_P0 = 0.0d0; RO_S0 = 0.0d0
  D_P0(1) = 0.20d0; RO_S0(1) = 2600.0d0
  RO_G0 = 0.0d0; MU_G0 = 0.0d0
  DES_CONTINUUM_COUPLED = .false.; MODEL_B = .false.
  DRAG_TYPE_ENUM = GIDASPOW
  DES_INTERP_SCHEME_ENUM = DES_INTERP_GARG
  DES_INTERP_ON = .false.; DES_INTERP_MEAN_FIELDS = .false.; DES_INTERP_WIDTH = UNDEFINED
  ! DEM namelist
  DES_MMAX = 1; MAX_PIP = 1; PIP = 1
  DES_COLL_MODEL_ENUM = LSD
  DES_INTG_METHOD = 'EULER'
  DTSOLID = UNDEFINED
  NEIGHBOR_SEARCH_N = 99; DES_NEIGHBOR_SEARCH = 4
  FACTOR_RLM = 1.2           ! default-REAL literals, as des_init_namelist.f sets them
  DESGRIDSEARCH_IMAX = 1; DESGRIDSEARCH_JMAX = 1; DESGRIDSEARCH_KMAX = 1
  REMOVE_ROGUE_PARTICLES = .true.; DES_EXPLICITLY_COUPLED = .false.
  KN = kn_in; KN_W = kn_in
  KT_FAC = 0.28571428571429; KT_W_FAC = 0.28571428571429
  DES_ETAT_FAC = UNDEFINED; DES_ETAT_W_FAC = UNDEFINED
  MEW = 0.0d0; MEW_W = 0.0d0
  DES_EN_INPUT = UNDEFINED; DES_ET_INPUT = UNDEFINED
  DES_EN_WALL_INPUT = UNDEFINED; DES_ET_WALL_INPUT = UNDEFINED
  DES_EN_INPUT(1) = en_in; DES_EN_WALL_INPUT(1) = en_in
  E_YOUNG = UNDEFINED; V_POISSON = UNDEFINED; EW_YOUNG = UNDEFINED; VW_POISSON = UNDEFINED
  MAX_RADIUS = 0.1d0; DES_EPG_CLIP = UNDEFINED
  ! particle arrays as MAKE_ARRAYS_DES leaves them
  allocate(DES_POS_NEW(MAX_PIP, 3), DES_VEL_NEW(MAX_PIP, 3), OMEGA_NEW(MAX_PIP, 3))
  allocate(DES_RADIUS(MAX_PIP), RO_SOL(MAX_PIP), PIJK(MAX_PIP, 5), PARTICLE_STATE(MAX_PIP))
  DES_POS_NEW(1, :) = [0.6d0, 0.5d0, 0.6d0]
  DES_VEL_NEW = 0.0d0; OMEGA_NEW = 0.0d0
  DES_RADIUS(1) = 0.1d0; RO_SOL(1) = 2600.0d0
  PIJK = 0; PIJK(1, 5) = 1
  PARTICLE_STATE = 1

  call INIT_PARTICLES_IN_CELL
  if (PIJK(1, 1) /= 2 .or. PIJK(1, 2) /= 2 .or. PIJK(1, 3) /= 2) error stop 'PIJK'
  do nn = 1, nsteps
     ! DES_TIME_STEP (des_time_march.f:293-460), uncoupled
     call CALC_FORCE_DEM
     call CALC_COORDINATION
     call CALC_DRAG_DES
     call CFNEWVALUES
     DO_NSEARCH = (nn == 1 .or. mod(nn, NEIGHBOR_SEARCH_N) == 0)
     if (DO_NSEARCH) then
        call mfixdem_fail(mfixdem_desgrid_pic(h), 'desgrid_pic')
        call NEIGHBOUR
     endif
     call mfixdem_fail(mfixdem_get_particles(h, DES_POS_NEW, DES_VEL_NEW, OMEGA_NEW, &
          PARTICLE_STATE, int(MAX_PIP, c_int64_t)), 'get_particles')
     write(*, '(i8, 2es25.16)') nn, DES_POS_NEW(1, 2), DES_VEL_NEW(1, 2)
  enddo
  call mfixdem_close()
end program udf_dem01
