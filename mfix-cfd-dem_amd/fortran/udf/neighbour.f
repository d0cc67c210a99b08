!----------------------------------------------------------------------!
! UDF override of model/des/neighbour.f (NEIGHBOUR_MOD): the neighbour  !
! list rebuild with the history carry of NEIGHBOUR (neighbour.f:21) on  !
! the GPU; DESGRID_PIC's binning is part of it.                        !
!----------------------------------------------------------------------!
MODULE NEIGHBOUR_MOD
  use mfixdem_c, only: mfixdem_neighbour
  use mfixdem_bridge, only: h, mfixdem_fail
  IMPLICIT NONE
CONTAINS
  SUBROUTINE NEIGHBOUR
    call mfixdem_fail(mfixdem_neighbour(h), 'neighbour')
  END SUBROUTINE NEIGHBOUR
END MODULE NEIGHBOUR_MOD
