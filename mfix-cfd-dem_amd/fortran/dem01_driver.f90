!----------------------------------------------------------------------!
! dem01_driver: the reference test case DEM01 (one sphere, d = 0.2 m,  !
! rho = 2600, dropped from y = 0.5 onto the south wall of a 1 m box,   !
! LSD, DES_EN = DES_EN_WALL) driven from Fortran through mfixdem_c.    !
! Usage: dem01_driver KN EN METHOD NSTEPS   (METHOD 0 Euler, 1 AB2)    !
! Prints "step time y vy" after every DES_TIME_STEP.                   !
!----------------------------------------------------------------------!
program dem01_driver
  use, intrinsic :: iso_c_binding
  use mfixdem_c
  implicit none
  type(mfixdem_params) :: p
  type(c_ptr) :: h
  real(c_double) :: pos(3), vel(3), om(3), rad(1), rho(1)
  integer(c_int32_t) :: ph(1)
  integer(c_int8_t) :: st(1)
  integer(c_int32_t) :: nn, nsteps, method
  real(c_double) :: kn, en, time
  character(len=64) :: arg

  kn = 1.0d4; en = 0.9d0; method = MFIXDEM_EULER; nsteps = 400
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg); read(arg, *) kn
  endif
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg); read(arg, *) en
  endif
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg); read(arg, *) method
  endif
  if (command_argument_count() >= 4) then
     call get_command_argument(4, arg); read(arg, *) nsteps
  endif

  call mfixdem_check(mfixdem_params_default(p), 'params_default')
  p%coll_model = MFIXDEM_LSD
  p%integrator = method
  p%nphases = 1
  p%max_pip = 1
  p%xlength = 1.0d0; p%ylength = 1.0d0; p%zlength = 1.0d0
  p%imax = 1; p%jmax = 1; p%kmax = 1
  p%desgridsearch_imax = 1; p%desgridsearch_jmax = 1; p%desgridsearch_kmax = 1
  p%neighbor_search_n = 99
  p%kn = kn; p%kn_w = kn
  p%mew = 0.0d0; p%mew_w = 0.0d0
  p%des_en_input(1) = en
  p%des_en_wall_input(1) = en
  p%d_p0(1) = 0.20d0
  p%ro_s0(1) = 2600.0d0
  p%max_radius = 0.1d0
  call mfixdem_check(mfixdem_create(p, h), 'create')

  pos = [0.6d0, 0.5d0, 0.6d0]
  vel = 0.0d0; om = 0.0d0
  rad(1) = 0.1d0; rho(1) = 2600.0d0; ph(1) = 1
  call mfixdem_check(mfixdem_set_particles(h, 1_c_int64_t, pos, vel, om, rad, rho, ph, &
       1_c_int64_t), 'set_particles')
  call mfixdem_check(mfixdem_init(h), 'init')
  time = 0.0d0
  do nn = 1, nsteps
     call mfixdem_check(mfixdem_step(h, nn), 'step')
     call mfixdem_check(mfixdem_get_particles(h, pos, vel, om, st, 1_c_int64_t), 'get')
     write(*, '(i8, 3es25.16)') nn, time, pos(2), vel(2)
  enddo
  call mfixdem_check(mfixdem_destroy(h), 'destroy')
end program dem01_driver
