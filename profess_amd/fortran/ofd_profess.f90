!------------------------------------------------------------------------------
! ISO_C_BINDING layer over libprofess_amd.so (include/ofd.h) and drop-in
! modules with the reference's names for the energy/potential path:
!   MODULE CalPotPlus   -> CalculatePotentialPlus(rho, optSqrt, potential, eTable)
!                          (CalPotential.f90:41)
!   MODULE RhoOptN      -> ProjNormRhoMinimization(energy, dir_type)
!                          (RhoOptN.f90:60), reading and writing SYS-style rhoR
!   MODULE OFD_Sys      -> the module state those routines read in the
!                          reference (rhoR, maxIter, pot_tol, tole, conv_check,
!                          rhoOutcome), plus the device context handle.
! A nonzero status from the library stops the program with the library's
! message, as the reference's QUIT does (Output.f90:1505-1549).
!------------------------------------------------------------------------------
MODULE ofd_c
  USE ISO_C_BINDING
  IMPLICIT NONE

  TYPE, BIND(C) :: ofd_grid_t
    INTEGER(C_INT) :: n1, n2, n3
    REAL(C_DOUBLE) :: cell(9)
  END TYPE ofd_grid_t

  TYPE, BIND(C) :: ofd_functional_t
    INTEGER(C_INT) :: kinetic
    REAL(C_DOUBLE) :: lambda, mu, alpha, beta, gamma, rho0, rhoS
    INTEGER(C_INT) :: precision
  END TYPE ofd_functional_t

  TYPE, BIND(C) :: ofd_opt_params_t
    INTEGER(C_INT) :: max_iter
    REAL(C_DOUBLE) :: pot_tol, tole
    INTEGER(C_INT) :: conv_check, max_line
    REAL(C_DOUBLE) :: mixed_switch
    INTEGER(C_INT) :: fused_dots
    INTEGER(C_INT) :: cg_alg
  END TYPE ofd_opt_params_t

  ! numSpin = 2 controls (CellInfo numSpin, XC_LDA lsda, RhoOptimizers
  ! fixedmag/nmag/tolm); see ofd_set_spin in include/ofd.h
  TYPE, BIND(C) :: ofd_spin_t
    INTEGER(C_INT) :: nspin, lsda, fixed_mag, nmag
    REAL(C_DOUBLE) :: tolm
  END TYPE ofd_spin_t

  TYPE, BIND(C) :: ofd_opt_stats_t
    INTEGER(C_INT) :: iterations, num_evals, newton_iters_total, fp32_evals
    REAL(C_DOUBLE) :: final_pot_norm, seconds
  END TYPE ofd_opt_stats_t

  INTERFACE
    FUNCTION ofd_create(grid, func, num_electrons, ctx) BIND(C, name="ofd_create")
      IMPORT :: C_INT, C_DOUBLE, C_PTR, ofd_grid_t, ofd_functional_t
      TYPE(ofd_grid_t), INTENT(IN) :: grid
      TYPE(ofd_functional_t), INTENT(IN) :: func
      REAL(C_DOUBLE), VALUE :: num_electrons
      TYPE(C_PTR), INTENT(OUT) :: ctx
      INTEGER(C_INT) :: ofd_create
    END FUNCTION
    SUBROUTINE ofd_destroy(ctx) BIND(C, name="ofd_destroy")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END SUBROUTINE
    FUNCTION ofd_last_error(ctx) BIND(C, name="ofd_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
      TYPE(C_PTR) :: ofd_last_error
    END FUNCTION
    FUNCTION ofd_set_ion_terms(ctx, vion, eii) BIND(C, name="ofd_set_ion_terms")
      IMPORT :: C_PTR, C_DOUBLE, C_INT
      TYPE(C_PTR), VALUE :: ctx, vion
      REAL(C_DOUBLE), VALUE :: eii
      INTEGER(C_INT) :: ofd_set_ion_terms
    END FUNCTION
    FUNCTION ofd_build_ion_potential(ctx, nspecies, type_start, frac, maxg, npts, poff, vals, order, eii) &
        BIND(C, name="ofd_build_ion_potential")
      IMPORT :: C_PTR, C_DOUBLE, C_INT
      TYPE(C_PTR), VALUE :: ctx, type_start, frac, maxg, npts, poff, vals
      INTEGER(C_INT), VALUE :: nspecies, order
      REAL(C_DOUBLE), VALUE :: eii
      INTEGER(C_INT) :: ofd_build_ion_potential
    END FUNCTION
    FUNCTION ofd_calpot(ctx, rho, opt_sqrt, pot, etable) BIND(C, name="ofd_calpot")
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: ctx, rho, pot, etable
      INTEGER(C_INT), VALUE :: opt_sqrt
      INTEGER(C_INT) :: ofd_calpot
    END FUNCTION
    FUNCTION ofd_rho_opt(ctx, rhoR, dir_type, params, energy, outcome, stats) BIND(C, name="ofd_rho_opt")
      IMPORT :: C_PTR, C_INT, ofd_opt_params_t, ofd_opt_stats_t
      TYPE(C_PTR), VALUE :: ctx, rhoR, energy
      INTEGER(C_INT), VALUE :: dir_type
      TYPE(ofd_opt_params_t), INTENT(IN) :: params
      INTEGER(C_INT), INTENT(OUT) :: outcome
      TYPE(ofd_opt_stats_t), INTENT(OUT) :: stats
      INTEGER(C_INT) :: ofd_rho_opt
    END FUNCTION
    FUNCTION ofd_ion_forces(ctx, rho, forces) BIND(C, name="ofd_ion_forces")
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: ctx, rho, forces
      INTEGER(C_INT) :: ofd_ion_forces
    END FUNCTION
    ! EwaldEnergy / EwaldForces (Ewald.f90:465-596, 1196-1310); forces and
    ! info5 may be C_NULL_PTR
    FUNCTION ofd_ewald(ctx, energy, forces, info5) BIND(C, name="ofd_ewald")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx, forces, info5
      REAL(C_DOUBLE), INTENT(OUT) :: energy
      INTEGER(C_INT) :: ofd_ewald
    END FUNCTION
    ! IonOptCG2::ConjugateGradient (IonOptCG2.f90:60-520); ion and stats are
    ! ofd_ion_params / ofd_ion_stats (C_PTR, may be C_NULL_PTR for ion)
    FUNCTION ofd_ion_relax(ctx, rhoR, frac, params, ion, energy, forces, energies, max_energies, stats) &
        BIND(C, name="ofd_ion_relax")
      IMPORT :: C_PTR, C_INT, ofd_opt_params_t
      TYPE(C_PTR), VALUE :: ctx, rhoR, frac, ion, energy, forces, energies, stats
      TYPE(ofd_opt_params_t), INTENT(IN) :: params
      INTEGER(C_INT), VALUE :: max_energies
      INTEGER(C_INT) :: ofd_ion_relax
    END FUNCTION
    FUNCTION ofd_set_spin(ctx, spin) BIND(C, name="ofd_set_spin")
      IMPORT :: C_PTR, C_INT, ofd_spin_t
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ofd_spin_t), INTENT(IN) :: spin
      INTEGER(C_INT) :: ofd_set_spin
    END FUNCTION
    FUNCTION ofd_get_magnetization(ctx, out3, mag_iters) BIND(C, name="ofd_get_magnetization")
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: out3(3)
      INTEGER(C_INT), INTENT(OUT) :: mag_iters
      INTEGER(C_INT) :: ofd_get_magnetization
    END FUNCTION
    FUNCTION c_strlen(s) BIND(C, name="strlen")
      IMPORT :: C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: s
      INTEGER(C_SIZE_T) :: c_strlen
    END FUNCTION
  END INTERFACE

CONTAINS

  ! Stop with the library's message (the reference's QUIT behaviour).
  SUBROUTINE ofd_check(ctx, status, where)
    TYPE(C_PTR), INTENT(IN) :: ctx
    INTEGER(C_INT), INTENT(IN) :: status
    CHARACTER(LEN=*), INTENT(IN) :: where
    TYPE(C_PTR) :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
    INTEGER :: n
    IF (status == 0) RETURN
    msg = ofd_last_error(ctx)
    n = INT(c_strlen(msg))
    CALL C_F_POINTER(msg, chars, [n])
    WRITE(*, '(A,A,A,*(A))') 'ofd: ', where, ': ', chars(1:n)
    STOP 1
  END SUBROUTINE ofd_check

END MODULE ofd_c


MODULE OFD_Sys
  USE ISO_C_BINDING
  IMPLICIT NONE
  TYPE(C_PTR), SAVE :: ofd_context = C_NULL_PTR
  INTEGER, SAVE :: n1G = 0, n2G = 0, n3G = 0
  ! SYS::rhoR (System.f90:79), rhoR(n1,n2,n3,numSpin)
  REAL(KIND=8), ALLOCATABLE, TARGET, SAVE :: rhoR(:,:,:,:)
  ! RhoOptimizers module variables (RhoOptimizers.f90:44-101)
  INTEGER, SAVE :: maxIter = 500
  INTEGER, SAVE :: rhoOutcome = -1
  REAL(KIND=8), SAVE :: pot_tol = 1.E-5
  REAL(KIND=8), SAVE :: tole = 1.E-6
  CHARACTER(LEN=500), SAVE :: conv_check = "ENE"
  INTEGER, SAVE :: numEnergyTotal = 0, numPotentialTotal = 0
  ! RhoDirCG::cg_alg (RhoDirCG.f90:26), 'PR' or 'HZ' (set by DENC)
  CHARACTER(LEN=500), SAVE :: cg_alg = 'HZ'
END MODULE OFD_Sys


MODULE CalPotPlus
  USE ISO_C_BINDING
  USE ofd_c
  USE OFD_Sys, ONLY: ofd_context
  IMPLICIT NONE
CONTAINS
  ! CalculatePotentialPlus(rho, optSqrt, potential, eTable) (CalPotential.f90:41):
  ! potential = dE/dphi (optSqrt) or dE/drho; eTable(9) in Hartree when present.
  SUBROUTINE CalculatePotentialPlus(rho, optSqrt, potential, eTable)
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(IN) :: rho
    LOGICAL, INTENT(IN) :: optSqrt
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(OUT) :: potential
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(OUT) :: eTable
    REAL(KIND=8), ALLOCATABLE, TARGET :: r(:,:,:,:), p(:,:,:,:)
    REAL(KIND=8), TARGET :: et(9)
    INTEGER(C_INT) :: st, os
    ! every spin channel (numSpin = SIZE(rho, 4), as set by ofd_set_spin)
    ALLOCATE(r(SIZE(rho,1), SIZE(rho,2), SIZE(rho,3), SIZE(rho,4)), &
             p(SIZE(rho,1), SIZE(rho,2), SIZE(rho,3), SIZE(rho,4)))
    r = rho              ! contiguous copy: callers pass expressions such as phi**2
    os = 0
    IF (optSqrt) os = 1
    IF (PRESENT(eTable)) THEN
      st = ofd_calpot(ofd_context, C_LOC(r), os, C_LOC(p), C_LOC(et))
      CALL ofd_check(ofd_context, st, 'CalculatePotentialPlus')
      eTable(1:9) = et
    ELSE
      st = ofd_calpot(ofd_context, C_LOC(r), os, C_LOC(p), C_NULL_PTR)
      CALL ofd_check(ofd_context, st, 'CalculatePotentialPlus')
    END IF
    potential = p
    DEALLOCATE(r, p)
  END SUBROUTINE CalculatePotentialPlus
END MODULE CalPotPlus


MODULE RhoOptN
  USE ISO_C_BINDING
  USE ofd_c
  USE OFD_Sys
  IMPLICIT NONE
CONTAINS
  ! ProjNormRhoMinimization(energy, dir_type) (RhoOptN.f90:60): optimizes
  ! OFD_Sys::rhoR in place with the truncated-Newton method (dir_type = 1),
  ! nonlinear CG (dir_type = 2, formula from cg_alg) or L-BFGS (dir_type = 3).
  SUBROUTINE ProjNormRhoMinimization(energy, dir_type)
    REAL(KIND=8), INTENT(OUT) :: energy(:)
    INTEGER, INTENT(INOUT) :: dir_type
    REAL(KIND=8), ALLOCATABLE, TARGET :: r(:,:,:,:)
    REAL(KIND=8), TARGET :: et(9)
    TYPE(ofd_opt_params_t) :: prm
    TYPE(ofd_opt_stats_t) :: stt
    INTEGER(C_INT) :: st, oc
    ALLOCATE(r(SIZE(rhoR,1), SIZE(rhoR,2), SIZE(rhoR,3), SIZE(rhoR,4)))
    r = rhoR
    prm%max_iter = maxIter
    prm%pot_tol = pot_tol
    prm%tole = tole
    prm%conv_check = 1
    IF (conv_check(1:3) == "POT") prm%conv_check = 2
    IF (conv_check(1:4) == "BOTH") prm%conv_check = 3
    prm%max_line = 200
    prm%mixed_switch = 0.d0
    prm%fused_dots = 0
    prm%cg_alg = 2
    IF (cg_alg(1:2) == 'PR') prm%cg_alg = 1
    st = ofd_rho_opt(ofd_context, C_LOC(r), INT(dir_type, C_INT), prm, C_LOC(et), oc, stt)
    CALL ofd_check(ofd_context, st, 'ProjNormRhoMinimization')
    rhoR = r
    energy(1:9) = et
    rhoOutcome = oc
    numEnergyTotal = stt%num_evals
    numPotentialTotal = stt%num_evals
    DEALLOCATE(r)
  END SUBROUTINE ProjNormRhoMinimization
END MODULE RhoOptN
