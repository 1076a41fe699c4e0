!------------------------------------------------------------------------------
! Fortran host exercising the drop-in modules on the reference's test/optIon
! input (4 Al, triclinic cell from displaced.ion, WT, 16^3 grid, recpot file
! al_HC.lda.recpot): one CalculatePotentialPlus at the uniform start density
! and a full ProjNormRhoMinimization. Prints the energies and returns a
! nonzero status when they miss the reference's recorded values
! (displaced.out:92-99: -219.9483598873 eV at step 0, -231.1425327622 eV
! converged; tolerance 1e-6 eV/atom on the converged value). Then the same
! cell with two spin channels (ofd_set_spin, fixed magnetization 0): at
! rho_up = rho_down = rho/2 every term but exchange-correlation equals the
! unpolarized one, and the spin solve reaches the unpolarized minimum.
! Usage: ofd_fortran_test <path to al_HC.lda.recpot>
!------------------------------------------------------------------------------
PROGRAM ofd_fortran_test
  USE ISO_C_BINDING
  USE ofd_c
  USE OFD_Sys
  USE CalPotPlus, ONLY: CalculatePotentialPlus
  USE RhoOptN, ONLY: ProjNormRhoMinimization
  IMPLICIT NONE
  REAL(KIND=8), PARAMETER :: bohr = 0.529177208607388d0
  REAL(KIND=8), PARAMETER :: pi = 3.141592653589793238462643d0
  REAL(KIND=8) :: h2ev, eii, vol
  REAL(KIND=8) :: latA(3,3)
  REAL(KIND=8), TARGET :: frac(3,4), maxg(1)
  INTEGER(C_INT), TARGET :: tstart(2), npts(1), poff(1)
  REAL(KIND=8), ALLOCATABLE, TARGET :: vals(:)
  REAL(KIND=8), ALLOCATABLE :: pot(:,:,:,:)
  REAL(KIND=8) :: energy(9), e0, ef, e_unpol(9), e_spin(9)
  REAL(KIND=8), ALLOCATABLE :: rho_conv(:,:,:)
  TYPE(ofd_spin_t) :: sp
  INTEGER :: k
  TYPE(ofd_grid_t) :: g
  TYPE(ofd_functional_t) :: f
  INTEGER(C_INT) :: st
  INTEGER :: dir_type, nerr
  CHARACTER(LEN=1024) :: path

  h2ev = 1.602176462d-19 * 1.d10 / (4.d0 * pi * (1.d0 / (4.d0 * pi * 1.d-7 * 299792458.d0**2)) * bohr)
  CALL GET_COMMAND_ARGUMENT(1, path)
  CALL read_recpot(TRIM(path), maxg(1), vals)

  ! displaced.ion: LATTICE_CART rows are lattice vectors (Angstrom)
  latA(:,1) = [3.869104452400000d0, 0.1d0, 0.0d0]
  latA(:,2) = [0.0d0, 3.969104452400000d0, 0.0d0]
  latA(:,3) = [0.0d0, 0.0d0, 4.0d0]
  frac(:,1) = [0.1d0, 0.0d0, 0.0d0]
  frac(:,2) = [0.5d0, 0.55d0, 0.99d0]
  frac(:,3) = [0.0d0, 0.5d0, 0.5d0]
  frac(:,4) = [0.505d0, 0.0d0, 0.5d0]
  g%n1 = 16; g%n2 = 16; g%n3 = 16
  g%cell = RESHAPE(latA / bohr, [9])
  f%kinetic = 4
  f%lambda = -100.d0; f%mu = -100.d0; f%alpha = -100.d0; f%beta = -100.d0
  f%gamma = -1.d0; f%rho0 = -1.d0; f%rhoS = -1.d0; f%precision = 0
  st = ofd_create(g, f, 12.d0, ofd_context)
  CALL ofd_check(ofd_context, st, 'ofd_create')
  eii = 0.d0                          ! replaced by the Ewald sum below
  tstart = [0, 4]
  npts(1) = SIZE(vals)
  poff(1) = 0
  st = ofd_build_ion_potential(ofd_context, 1, C_LOC(tstart), C_LOC(frac), C_LOC(maxg), C_LOC(npts), &
                               C_LOC(poff), C_LOC(vals), 10, eii)
  CALL ofd_check(ofd_context, st, 'ofd_build_ion_potential')
  ! ionIonEnergy = EwaldEnergy(...) as RefreshIonTerms sets it (RefreshIons.f90:196)
  st = ofd_ewald(ofd_context, eii, C_NULL_PTR, C_NULL_PTR)
  CALL ofd_check(ofd_context, st, 'ofd_ewald')
  WRITE(*, '(A,ES20.12,A)') ' ion-ion energy     ', eii * h2ev, ' eV'

  vol = ABS(det3(latA / bohr))
  n1G = 16; n2G = 16; n3G = 16
  ALLOCATE(rhoR(16,16,16,1), pot(16,16,16,1))
  rhoR = 12.d0 / vol                  ! uniform start (System.f90:320)
  CALL CalculatePotentialPlus(rhoR, .TRUE., pot, energy)
  e0 = energy(1) * h2ev
  WRITE(*, '(A,F20.10,A)') ' step 0 energy      ', e0, ' eV'
  dir_type = 1
  CALL ProjNormRhoMinimization(energy, dir_type)
  ef = energy(1) * h2ev
  WRITE(*, '(A,F20.10,A,I3,A,I4)') ' converged energy   ', ef, ' eV  outcome', rhoOutcome, &
                                   '  evaluations', numEnergyTotal
  WRITE(*, '(A,9ES20.12)') ' eTable (Ha)', energy
  nerr = 0
  IF (ABS(e0 - (-2.199483598873d2)) > 1.d-9 * 219.95d0) nerr = nerr + 1
  IF (ABS(ef - (-2.311425327622d2)) / 4.d0 > 1.d-6) nerr = nerr + 1
  IF (rhoOutcome /= 0) nerr = nerr + 1
  IF (ABS(eii * h2ev - (-2.986846473280d2)) > 1.d-10 * 298.7d0) nerr = nerr + 1   ! displaced.out:119

  ! numSpin = 2 at the converged density, split evenly
  ALLOCATE(rho_conv(16,16,16))
  rho_conv = rhoR(:,:,:,1)
  CALL CalculatePotentialPlus(rhoR, .TRUE., pot, e_unpol)
  sp%nspin = 2; sp%lsda = 1; sp%fixed_mag = 1; sp%nmag = -1; sp%tolm = -1.d0
  st = ofd_set_spin(ofd_context, sp)
  CALL ofd_check(ofd_context, st, 'ofd_set_spin')
  DEALLOCATE(rhoR, pot)
  ALLOCATE(rhoR(16,16,16,2), pot(16,16,16,2))
  rhoR(:,:,:,1) = 0.5d0 * rho_conv
  rhoR(:,:,:,2) = 0.5d0 * rho_conv
  CALL CalculatePotentialPlus(rhoR, .TRUE., pot, e_spin)
  WRITE(*, '(A,9ES20.12)') ' spin eTable (Ha)', e_spin
  DO k = 2, 9        ! kinetic, ion-electron, Hartree, ion-ion, TF, vW, WT (Fortran slots)
    IF (k == 5) CYCLE
    IF (ABS(e_spin(k) - e_unpol(k)) > 1.d-12 * ABS(e_unpol(k)) + 1.d-14) nerr = nerr + 1
  END DO
  ! LSDAPotPZ's rounded exchange constant (-0.7385588) against LDA's
  IF (ABS(e_spin(5) - e_unpol(5)) > 1.d-7 * ABS(e_unpol(5))) nerr = nerr + 1
  CALL ProjNormRhoMinimization(energy, dir_type)
  WRITE(*, '(A,F20.10,A,I3)') ' spin converged     ', energy(1) * h2ev, ' eV  outcome', rhoOutcome
  IF (rhoOutcome /= 0) nerr = nerr + 1
  IF (ABS(energy(1) * h2ev - ef) / 4.d0 > 1.d-5) nerr = nerr + 1
  CALL ofd_destroy(ofd_context)
  IF (nerr /= 0) THEN
    WRITE(*, '(A,I2,A)') ' FAILED (', nerr, ' checks)'
    STOP 1
  END IF
  WRITE(*, '(A)') ' PASSED'

CONTAINS

  FUNCTION det3(a) RESULT(d)
    REAL(KIND=8), INTENT(IN) :: a(3,3)
    REAL(KIND=8) :: d
    d = a(1,1)*(a(2,2)*a(3,3)-a(2,3)*a(3,2)) - a(1,2)*(a(2,1)*a(3,3)-a(2,3)*a(3,1)) &
      + a(1,3)*(a(2,1)*a(3,2)-a(2,2)*a(3,1))
  END FUNCTION det3

  ! .recpot reading rules of ReadPseudo (ReadIonFile.f90:830-870): skip to
  ! END COMMENT, one line, maxG (1/A), then three values per line up to the
  ! "1000" terminator.
  SUBROUTINE read_recpot(fname, mg, v)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    REAL(KIND=8), INTENT(OUT) :: mg
    REAL(KIND=8), ALLOCATABLE, INTENT(OUT) :: v(:)
    CHARACTER(LEN=256) :: line
    INTEGER :: ios, n, u
    OPEN(NEWUNIT=u, FILE=fname, STATUS='old', ACTION='read')
    line = ''
    DO WHILE (INDEX(line, 'END COMMENT') == 0)
      READ(u, '(A)') line
    END DO
    READ(u, '(A)') line
    READ(u, *) mg
    n = 0
    DO
      READ(u, '(A)', IOSTAT=ios) line
      IF (ios /= 0) EXIT
      IF (ADJUSTL(line(1:8)) == '1000') EXIT
      n = n + 3
    END DO
    REWIND(u)
    line = ''
    DO WHILE (INDEX(line, 'END COMMENT') == 0)
      READ(u, '(A)') line
    END DO
    READ(u, '(A)') line
    READ(u, '(A)') line
    ALLOCATE(v(n))
    READ(u, *) v
    CLOSE(u)
  END SUBROUTINE read_recpot

END PROGRAM ofd_fortran_test
