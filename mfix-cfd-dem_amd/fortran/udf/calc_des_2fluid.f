!----------------------------------------------------------------------!
! UDF override of model/des/calc_des_2fluid.f (CALC_DES_2FLUID_MOD):    !
! the gas side of the drag for the fluid solver (calc_des_2fluid.f:25: !
! DRAG_GS_GAS0 for GARG_2012, DRAG_GS_GAS1 otherwise) on the GPU.  The !
! new gas field goes to the GPU first.  Interphase heat and species    !
! transfer are outside the library.                                    !
!----------------------------------------------------------------------!
MODULE CALC_DES_2FLUID_MOD
  use mfixdem_c, only: mfixdem_drag_gs_gas, mfixdem_get_gas_drag
  use mfixdem_bridge, only: h, mfixdem_fail, mfixdem_push_gas
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CALC_DES_2FLUID
    use run, only: DES_CONTINUUM_COUPLED
    use discretelement, only: F_GDS, DRAG_AM, DRAG_BM
    IF (.NOT. DES_CONTINUUM_COUPLED) RETURN
    call mfixdem_push_gas()
    call mfixdem_fail(mfixdem_drag_gs_gas(h), 'drag_gs_gas')
    call mfixdem_fail(mfixdem_get_gas_drag(h, F_GDS, DRAG_AM, DRAG_BM), 'get_gas_drag')
  END SUBROUTINE CALC_DES_2FLUID

  SUBROUTINE DES_2FLUID_CONV(S_P, S_C)
    DOUBLE PRECISION, INTENT(INOUT) :: S_P(*), S_C(*)
    call mfixdem_fail(9, 'interphase heat transfer')
  END SUBROUTINE DES_2FLUID_CONV

  SUBROUTINE DES_2FLUID_RXNS
    call mfixdem_fail(9, 'interphase species transfer')
  END SUBROUTINE DES_2FLUID_RXNS
END MODULE CALC_DES_2FLUID_MOD
