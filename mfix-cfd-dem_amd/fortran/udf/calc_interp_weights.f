!----------------------------------------------------------------------!
! UDF override of model/des/calc_interp_weights.f                       !
! (CALC_INTERP_WEIGHTS_MOD).  The library evaluates the filter weights  !
! inside its drag and deposition kernels, where they are used, so the  !
! routines only hand the step to it; FILTER_CELL / FILTER_WEIGHT are   !
! not filled on the host.                                              !
!----------------------------------------------------------------------!
MODULE CALC_INTERP_WEIGHTS_MOD
  use mfixdem_c, only: mfixdem_calc_interp_weights
  use mfixdem_bridge, only: h, mfixdem_fail
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CALC_INTERP_WEIGHTS
    call mfixdem_fail(mfixdem_calc_interp_weights(h), 'calc_interp_weights')
  END SUBROUTINE CALC_INTERP_WEIGHTS
  SUBROUTINE CALC_INTERP_WEIGHTS1
    call CALC_INTERP_WEIGHTS
  END SUBROUTINE CALC_INTERP_WEIGHTS1
  SUBROUTINE CALC_INTERP_WEIGHTS2
    call CALC_INTERP_WEIGHTS
  END SUBROUTINE CALC_INTERP_WEIGHTS2
  SUBROUTINE CALC_INTERP_WEIGHTS3
    call CALC_INTERP_WEIGHTS
  END SUBROUTINE CALC_INTERP_WEIGHTS3
END MODULE CALC_INTERP_WEIGHTS_MOD
