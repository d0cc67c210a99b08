!----------------------------------------------------------------------!
! UDF override of model/des/calc_epg_des.f (CALC_EPG_DES_MOD): EP_G is  !
! formed on the GPU by the deposition itself (mfixdem_comp_mean_fields !
! ends with it, calc_epg_des.f:19), so this routine has nothing left.  !
!----------------------------------------------------------------------!
MODULE CALC_EPG_DES_MOD
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CALC_EPG_DES
  END SUBROUTINE CALC_EPG_DES
END MODULE CALC_EPG_DES_MOD
