!----------------------------------------------------------------------!
! UDF override of model/des/comp_mean_fields.f (COMP_MEAN_FIELDS_MOD):  !
! the solids deposition and EP_G (comp_mean_fields.f:22, with          !
! CALC_EPG_DES) on the GPU; EP_G and ROP_S come back to the fluid      !
! arrays.                                                              !
!----------------------------------------------------------------------!
MODULE COMP_MEAN_FIELDS_MOD
  use mfixdem_c, only: mfixdem_comp_mean_fields, mfixdem_get_gas
  use mfixdem_bridge, only: h, mfixdem_fail
  IMPLICIT NONE
CONTAINS
  SUBROUTINE COMP_MEAN_FIELDS
    use physprop, only: MMAX
    use fldvar, only: EP_G, ROP_S
    call mfixdem_fail(mfixdem_comp_mean_fields(h), 'comp_mean_fields')
    call mfixdem_fail(mfixdem_get_gas(h, EP_G, ROP_S(1, MMAX + 1)), 'get_gas')
  END SUBROUTINE COMP_MEAN_FIELDS
END MODULE COMP_MEAN_FIELDS_MOD
