! ref_probe.f90 -- test harness over the reference's own Fortran modules.
!
! TEST INFRASTRUCTURE ONLY. It is linked against the reference's src/*.f90 modules
! (built in place by oracle/build_ref.sh into oracle/_ref/) and evaluates reference
! functions on inputs read from stdin, printing the raw IEEE bit patterns so that
! tests/golden/gen_golden.py can store them as fixtures.  It is never part of the
! product path.
!
! Commands (one per line on stdin):
!   const
!       -> bit patterns of PI, sqln2, C2, dopplerCONST, sqrt(pi), sqrt(1.4), log(2.)
!          (src/Constants.f90:7-22)
!   voigt <X> <gD> <gL> <S> <mol> <chi> <T>
!       -> voigt(X) with dopHWHM=gD, lorHWHM=gL, lineIntensity=S (src/Shapes.f90:22-133)
!          chi: 0 none, 1 tonkov, 2 pollack, 3 perrin (src/ChiFactors.f90:18-122)
!   lorshape <X> <gL> <S> <mol> <chi> <T>
!       -> lorentzShape(X) (src/Shapes.f90:135-141)
!   chi <id> <X> <mol> <T>
!       -> chi factor value
!   spec <nu> <S> <gair> <gself> <E> <n> <molIso> <delta> <P> <T> <rho>
!       -> pSelf (app/main.f90:92), shifted nu (Spectroscopy.f90:38), dopplerHWHM (:52),
!          lorentzHWHM (:69), TIPSofT (:91), intensityOfT (:118)
!   cascade <file>
!       -> reads 3*10230 + 20481 FP32 bit patterns (hex, one per line: for k = 0..9 the
!          P, mid, L arrays of grid k, then RK) into the grid arrays, runs
!          cascadeInterpolation (src/Grids.f90:104-139) and prints RK(1..20481)
!   wiso
!       -> bit patterns of WISO(1..124) (src/MolarMasses.f90:140-146)
!   tips
!       -> bit patterns of TIPS(74,501) as read by the reference's reader, row by row
!   quit
program ref_probe
    use Constants
    use Interfaces
    use Atmosphere
    use MolarMasses
    use ChiFactors
    use Spectroscopy
    use Shapes
    use Grids
    implicit none

    character(len=32) :: cmd
    character(len=512) :: line
    integer :: qios, chiId, mol, iso
    real(kind=DP) :: X, nuIn
    real :: gD, gL, qS, qT, qP, qrho, gair, gself, qE, qn, qdelta
    real :: r
    real(kind=DP) :: shifted
    integer :: i
    character(len=400) :: fname

    call readTIPS_local()

    do
        read(*, '(A)', iostat=qios) line
        if (qios /= 0) exit
        read(line, *, iostat=qios) cmd
        if (qios /= 0) cycle
        select case (trim(cmd))
        case ('quit')
            exit
        case ('const')
            write(*, '(7(Z8.8,1X))') transfer(PI, 0), transfer(sqln2, 0), transfer(C2, 0), &
                transfer(dopplerCONST, 0), transfer(sqrt(pi), 0), transfer(sqrt(1.4), 0), &
                transfer(log(2.), 0)
        case ('tips')
            do i = 1, size(TIPS, 1)
                write(*, '(501(Z8.8,1X))') (transfer(TIPS(i, chiId), 0), chiId = 1, size(TIPS, 2))
            end do
        case ('wiso')
            do i = 1, size(WISO)
                write(*, '(Z8.8)') transfer(WISO(i), 0)
            end do
        case ('voigt')
            read(line, *) cmd, X, gD, gL, qS, mol, chiId, qT
            call setChi(chiId)
            moleculeIntCode = mol
            temperature = qT
            dopHWHM = gD
            lorHWHM = gL
            lineIntensity = qS
            r = voigt(X)
            write(*, '(Z8.8)') transfer(r, 0)
        case ('lorshape')
            read(line, *) cmd, X, gL, qS, mol, chiId, qT
            call setChi(chiId)
            moleculeIntCode = mol
            temperature = qT
            lorHWHM = gL
            lineIntensity = qS
            r = lorentzShape(X)
            write(*, '(Z8.8)') transfer(r, 0)
        case ('chi')
            read(line, *) cmd, chiId, X, mol, qT
            call setChi(chiId)
            temperature = qT
            r = chiFactorFuncPtr(X, mol)
            write(*, '(Z8.8)') transfer(r, 0)
        case ('spec')
            read(line, *) cmd, nuIn, qS, gair, gself, qE, qn, iso, qdelta, qP, qT, qrho
            lineWV = nuIn
            refLineIntensity = qS
            gammaForeign = gair
            gammaSelf = gself
            lineLowerState = qE
            foreignTempCoeff = qn
            jointMolIso = iso
            deltaForeign = qdelta
            pressure = qP
            temperature = qT
            density = qrho
            pSelf = density * 10. * BOLsi * temperature / standardAtmosphericPressure
            isotopeNum = jointMolIso / 100
            molarMass = WISO(isotopeNum)
            shifted = shiftedLinePosition(lineWV, pressure)
            dopHWHM = dopplerHWHM(shifted, temperature, molarMass)
            lorHWHM = lorentzHWHM(pressureParameter=pressure, partialPressureParameter=pSelf, &
                                  temperatureParameter=temperature)
            lineIntensity = intensityOfT(temperature)
            write(*, '(Z8.8,1X,Z16.16,4(1X,Z8.8))') transfer(pSelf, 0), transfer(shifted, 0_8), &
                transfer(dopHWHM, 0), transfer(lorHWHM, 0), transfer(TIPSofT(temperature), 0), &
                transfer(lineIntensity, 0)
        case ('cascade')
            read(line, *) cmd, fname
            call loadGrids(trim(fname))
            call cascadeInterpolation()
            do i = 1, NT
                write(*, '(Z8.8)') transfer(RK(i), 0)
            end do
        end select
    end do

contains

    subroutine setChi(id)
        integer, intent(in) :: id
        select case (id)
        case (1)
            chiFactorFuncPtr => tonkov
        case (2)
            chiFactorFuncPtr => pollack
        case (3)
            chiFactorFuncPtr => perrin
        case default
            chiFactorFuncPtr => noneChi
        end select
    end subroutine setChi

    subroutine readTIPS_local()
        ! same file layout as the reference's readTIPS (src/EnvironmentSetup.f90:169-184)
        integer :: nIso, nTemp, a, b
        open(unit=5467, file='data/TIPS/TIPS.dat', status='old', action='read')
        read(5467, *) nIso, nTemp
        allocate(TIPS(nIso, nTemp))
        do a = 1, nIso
            read(5467, *) (TIPS(a, b), b = 1, nTemp)
        end do
        close(5467)
    end subroutine readTIPS_local

    subroutine loadGrids(fn)
        character(len=*), intent(in) :: fn
        open(unit=901, file=fn, status='old', action='read')
        call load3(RK0P, RK0, RK0L)
        call load3(RK1P, RK1, RK1L)
        call load3(RK2P, RK2, RK2L)
        call load3(RK3P, RK3, RK3L)
        call load3(RK4P, RK4, RK4L)
        call load3(RK5P, RK5, RK5L)
        call load3(RK6P, RK6, RK6L)
        call load3(RK7P, RK7, RK7L)
        call load3(RK8P, RK8, RK8L)
        call load3(RK9P, RK9, RK9L)
        call load1(RK)
        close(901)
    end subroutine loadGrids

    subroutine load3(a, b, c)
        real, intent(inout) :: a(:), b(:), c(:)
        call load1(a)
        call load1(b)
        call load1(c)
    end subroutine load3

    subroutine load1(a)
        real, intent(inout) :: a(:)
        integer :: k, bits
        do k = 1, size(a)
            read(901, '(Z8)') bits
            a(k) = transfer(bits, 1.0)
        end do
    end subroutine load1

end program ref_probe
