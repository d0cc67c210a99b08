!----------------------------------------------------------------------!
! UDF override of model/des/calc_drag_des.f (CALC_DRAG_DES_MOD): the    !
! gas-solids drag on each particle (calc_drag_des.f:13) on the GPU.    !
!----------------------------------------------------------------------!
MODULE CALC_DRAG_DES_MOD
  use mfixdem_c, only: mfixdem_calc_drag_des
  use mfixdem_bridge, only: h, mfixdem_fail
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CALC_DRAG_DES
    call mfixdem_fail(mfixdem_calc_drag_des(h), 'calc_drag_des')
  END SUBROUTINE CALC_DRAG_DES
END MODULE CALC_DRAG_DES_MOD
