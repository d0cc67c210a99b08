!----------------------------------------------------------------------!
! Stand-ins for the MFiX modules the UDF overrides USE: the module and  !
! variable names, kinds and shapes of model/param_mod.f, param1_mod.f, !
! constant_mod.f, geometry_mod.f, physprop_mod.f, run_mod.f,           !
! fldvar_mod.f, des/derived_types_mod.f, des/particle_filter_mod.f and !
! des/discretelement_mod.f, with nothing else.  Test infrastructure:   !
! the overrides compile against these as they would against MFiX, and !
! a driver fills them to run a case.                                   !
!----------------------------------------------------------------------!
module param
  implicit none
  integer, parameter :: DIM_M = 10, DIM_I = 1024, DIM_J = 1024, DIM_K = 1024
end module param

module param1
  implicit none
  double precision, parameter :: UNDEFINED = 9.87654321d31, ZERO = 0.0d0
  integer, parameter :: UNDEFINED_I = 987654321
end module param1

module constant
  implicit none
  double precision :: GRAVITY, GRAVITY_X, GRAVITY_Y, GRAVITY_Z
end module constant

module geometry
  use param, only: DIM_I, DIM_J, DIM_K
  implicit none
  double precision :: XLENGTH, YLENGTH, ZLENGTH
  double precision :: X_MIN, X_MAX, Y_MIN, Y_MAX, Z_MIN, Z_MAX
  integer :: IMIN1, JMIN1, KMIN1
  integer :: IMAX, JMAX, KMAX
  double precision :: DX(0:DIM_I+2), DY(0:DIM_J+2), DZ(0:DIM_K+2)
end module geometry

module physprop
  use param, only: DIM_M
  implicit none
  integer :: MMAX
  double precision :: D_p0(DIM_M), RO_s0(DIM_M)
  double precision :: RO_g0, MU_g0
end module physprop

module run
  implicit none
  logical :: DES_CONTINUUM_COUPLED
  logical :: Model_B
end module run

module fldvar
  implicit none
  double precision, dimension(:), allocatable :: EP_g, P_g, U_g, V_g, W_g
  double precision, dimension(:, :), allocatable :: ROP_s
end module fldvar

module derived_types
  implicit none
  integer :: DRAG_TYPE_ENUM
  integer, parameter :: GIDASPOW = 1, WEN_YU = 5
end module derived_types

module particle_filter
  implicit none
  logical :: DES_INTERP_ON, DES_INTERP_MEAN_FIELDS
  double precision :: DES_INTERP_WIDTH
  integer :: DES_INTERP_SCHEME_ENUM
  integer, parameter :: DES_INTERP_NONE = 0, DES_INTERP_GARG = 1, DES_INTERP_DPVM = 2
  integer, parameter :: DES_INTERP_GAUSS = 3, DES_INTERP_LHAT = 4
end module particle_filter

module discretelement
  use param, only: DIM_M
  implicit none
  integer, parameter :: DIM_M_TRI = 55
  integer :: PIP, MAX_PIP, DES_MMAX
  integer :: DES_COLL_MODEL_ENUM
  integer, parameter :: HERTZIAN = 0, LSD = 1
  character(len=64) :: DES_INTG_METHOD
  double precision :: DTSOLID
  integer :: NEIGHBOR_SEARCH_N, DES_NEIGHBOR_SEARCH
  logical :: DO_NSEARCH, REMOVE_ROGUE_PARTICLES, DES_EXPLICITLY_COUPLED
  double precision :: FACTOR_RLM
  integer :: DESGRIDSEARCH_IMAX, DESGRIDSEARCH_JMAX, DESGRIDSEARCH_KMAX
  double precision :: KN, KN_W, KT, KT_W, KT_FAC, KT_W_FAC
  double precision :: DES_ETAT_FAC, DES_ETAT_W_FAC
  double precision :: MEW, MEW_W
  double precision :: DES_EN_INPUT(DIM_M_TRI), DES_ET_INPUT(DIM_M_TRI)
  double precision :: DES_EN_WALL_INPUT(DIM_M), DES_ET_WALL_INPUT(DIM_M)
  double precision :: E_YOUNG(DIM_M), Ew_YOUNG, V_POISSON(DIM_M), Vw_POISSON
  double precision :: MIN_RADIUS, MAX_RADIUS
  double precision :: DES_EPG_CLIP
  integer(kind=1), dimension(:), allocatable :: PARTICLE_STATE
  double precision, dimension(:), allocatable :: DES_RADIUS, RO_Sol
  double precision, dimension(:,:), allocatable :: DES_POS_NEW, DES_VEL_NEW, OMEGA_NEW
  integer, dimension(:,:), allocatable :: PIJK
  integer, dimension(:), allocatable :: NEIGHBOR_INDEX, NEIGHBORS
  double precision, dimension(:,:), allocatable :: PFT_NEIGHBOR
  double precision, dimension(:), allocatable :: F_GDS, DRAG_AM
  double precision, dimension(:,:), allocatable :: DRAG_BM, P_FORCE, DRAG_FC
end module discretelement
