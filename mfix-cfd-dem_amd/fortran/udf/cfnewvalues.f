!----------------------------------------------------------------------!
! UDF override of model/des/cfnewvalues.f (CFNEWVALUES_MOD): the Euler  !
! / Adams-Bashforth update (cfnewvalues.f:641) on the GPU.  Glued-     !
! sphere, super-quadric and prescribed rigid-body motion are outside   !
! the library and stop the run if a case reaches them.                 !
!----------------------------------------------------------------------!
MODULE CFNEWVALUES_MOD
  use mfixdem_c, only: mfixdem_cfnewvalues
  use mfixdem_bridge, only: h, mfixdem_fail
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CFNEWVALUES
    call mfixdem_fail(mfixdem_cfnewvalues(h), 'cfnewvalues')
  END SUBROUTINE CFNEWVALUES

  ! No prescribed particle motion in the library's cases.
  SUBROUTINE DEM_SET_RIGID_MOTION
  END SUBROUTINE DEM_SET_RIGID_MOTION

  SUBROUTINE UPDATE_GSP_PARTICLE_STATE
    call mfixdem_fail(9, 'glued-sphere particles')
  END SUBROUTINE UPDATE_GSP_PARTICLE_STATE
  SUBROUTINE GSP_CFNEWVALUES_vhalfxfull
    call mfixdem_fail(9, 'glued-sphere particles')
  END SUBROUTINE GSP_CFNEWVALUES_vhalfxfull
  SUBROUTINE GSP_CFNEWVALUES_vfull
    call mfixdem_fail(9, 'glued-sphere particles')
  END SUBROUTINE GSP_CFNEWVALUES_vfull
  SUBROUTINE GSP_CFNEWVALUES_vhalfxfull_implicit
    call mfixdem_fail(9, 'glued-sphere particles')
  END SUBROUTINE GSP_CFNEWVALUES_vhalfxfull_implicit
  SUBROUTINE GSP_CFNEWVALUES_vfull_implicit
    call mfixdem_fail(9, 'glued-sphere particles')
  END SUBROUTINE GSP_CFNEWVALUES_vfull_implicit
  SUBROUTINE SuperDEM_CFNEWVALUES
    call mfixdem_fail(9, 'super-quadric particles')
  END SUBROUTINE SuperDEM_CFNEWVALUES
END MODULE CFNEWVALUES_MOD
