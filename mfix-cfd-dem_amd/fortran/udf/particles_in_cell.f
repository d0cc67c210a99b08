!----------------------------------------------------------------------!
! UDF override of model/des/particles_in_cell.f (PARTICLES_IN_CELL_MOD  !
! and the PIC_SEARCH utility): the fluid cell of each particle         !
! (particles_in_cell.f:20) on the GPU.  INIT_PARTICLES_IN_CELL, which  !
! MAKE_ARRAYS_DES calls once the particles are set up                  !
! (make_arrays_des.f:343), is where the case's handle is opened: the   !
! particles go to the GPU there, are located in the fluid grid (PIJK   !
! of every particle, coupled or not, as INIT_PARTICLES_IN_CELL sets    !
! it) and PIJK comes back.                                             !
!----------------------------------------------------------------------!
MODULE PARTICLES_IN_CELL_MOD
  use mfixdem_c, only: mfixdem_particles_in_cell
  use mfixdem_bridge, only: h, mfixdem_fail, mfixdem_open, mfixdem_pull
  IMPLICIT NONE
CONTAINS
  SUBROUTINE PARTICLES_IN_CELL
    call mfixdem_fail(mfixdem_particles_in_cell(h), 'particles_in_cell')
  END SUBROUTINE PARTICLES_IN_CELL

  SUBROUTINE INIT_PARTICLES_IN_CELL
    call mfixdem_open()
    call mfixdem_fail(mfixdem_particles_in_cell(h), 'particles_in_cell')
    call mfixdem_pull()
  END SUBROUTINE INIT_PARTICLES_IN_CELL
END MODULE PARTICLES_IN_CELL_MOD

! The cell IDX in lSTART..lEND with ENT_POS(IDX-1) <= lPOS < ENT_POS(IDX),
! lEND+1 when there is none (particles_in_cell.f:347).
SUBROUTINE PIC_SEARCH(IDX, lPOS, ENT_POS, lDIMN, lSTART, lEND)
  IMPLICIT NONE
  INTEGER, INTENT(OUT) :: IDX
  DOUBLE PRECISION, INTENT(IN) :: lPOS
  INTEGER, INTENT(IN) :: lDIMN
  DOUBLE PRECISION, INTENT(IN) :: ENT_POS(0:lDIMN)
  INTEGER, INTENT(IN) :: lSTART, lEND
  INTEGER :: C
  IDX = lEND + 1
  DO C = lSTART, lEND
     IF (ENT_POS(C-1) <= lPOS .AND. lPOS < ENT_POS(C)) THEN
        IDX = C
        RETURN
     ENDIF
  ENDDO
END SUBROUTINE PIC_SEARCH
