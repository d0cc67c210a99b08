!----------------------------------------------------------------------!
! UDF override of model/des/calc_pg_grad.f (CALC_PG_GRAD_MOD): P_FORCE  !
! = -grad P_G (calc_pg_grad.f:20) on the GPU, once per fluid step;     !
! P_FORCE comes back for the fluid side.                               !
!----------------------------------------------------------------------!
MODULE CALC_PG_GRAD_MOD
  use mfixdem_bridge, only: mfixdem_push_pressure
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CALC_PG_GRAD
    call mfixdem_push_pressure()
  END SUBROUTINE CALC_PG_GRAD
END MODULE CALC_PG_GRAD_MOD
