!----------------------------------------------------------------------!
! UDF override of model/des/calc_force_dem.f (CALC_FORCE_DEM_MOD):      !
! particle-particle and particle-wall contact forces on the GPU        !
! (calc_force_dem.f:18, calc_collision_wall_mod.f:54).                 !
!----------------------------------------------------------------------!
MODULE CALC_FORCE_DEM_MOD
  use mfixdem_c, only: mfixdem_calc_force_dem
  use mfixdem_bridge, only: h, mfixdem_fail
  IMPLICIT NONE
CONTAINS

  SUBROUTINE CALC_FORCE_DEM
    call mfixdem_fail(mfixdem_calc_force_dem(h), 'calc_force_dem')
  END SUBROUTINE CALC_FORCE_DEM

  ! Coordination numbers are an output diagnostic, not part of the step.
  SUBROUTINE CALC_COORDINATION
  END SUBROUTINE CALC_COORDINATION

  ! Overlap report of the reference's debug path: nothing to report.
  SUBROUTINE PRINT_EXCESS_OVERLAP
  END SUBROUTINE PRINT_EXCESS_OVERLAP

  SUBROUTINE gsp_gsp_collison(index_ll, index_i, index_CC, VRN, VSLIP, NORM, overlapn, &
       dist2I, dist2LL)
    INTEGER, INTENT(IN) :: index_ll, index_i, index_CC
    DOUBLE PRECISION :: VRN, VSLIP(3), NORM(3), overlapn, dist2I(3), dist2LL(3)
    call mfixdem_fail(9, 'glued-sphere collisions')
  END SUBROUTINE gsp_gsp_collison

END MODULE CALC_FORCE_DEM_MOD
