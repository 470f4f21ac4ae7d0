! ref_driver.f90 -- level/subinterval driver over the reference's own LBL modules.
!
! TEST INFRASTRUCTURE ONLY (golden vectors and the timed CPU baseline).  It is linked
! against the reference's src/*.f90 modules, built in place by oracle/build_ref.sh into
! oracle/_ref/.  The loops below restate app/main.f90:81-180 (levels, subintervals,
! record numbers, window start carry) and app/main.f90:196-248 (starting line), but take
! explicit file paths and a chi-factor selector, because the reference binds the chi
! factor at compile time (src/ChiFactors.f90:11-16) and hard-codes its data paths.
! With chi = 0 its records are byte-identical to the reference's own `marfa` program,
! which tests/golden/gen_golden.py checks.
!
! usage: ref_driver startWV endWV dbFile cutOff ACS|VAC atmFile chiId molCode outDir
!        [maxLevels [shapeId]]      shapeId 0 voigt (default), 1 lorentzShape
! TIPS is read from data/TIPS/TIPS.dat relative to the working directory, as the
! reference does (src/EnvironmentSetup.f90:238).
program ref_driver
    use Constants
    use Atmosphere
    use Spectroscopy
    use ChiFactors
    use Shapes
    use LBL
    use Grids
    implicit none

    integer, parameter :: outUnit = 47
    character(len=400) :: arg, dbFile, atmFile, outDir
    character(len=8) :: target
    character(len=3) :: levelLabel
    integer :: chiId, maxLevels, lvl, recNum, startIdx, lineIdx, shapeId
    real(kind=DP) :: startLineWV, capWV, t0, t1
    logical :: first

    call get_command_argument(1, arg); read(arg, *) startWV
    call get_command_argument(2, arg); read(arg, *) endWV
    call get_command_argument(3, dbFile)
    call get_command_argument(4, arg); read(arg, *) cutOff
    call get_command_argument(5, target)
    call get_command_argument(6, atmFile)
    call get_command_argument(7, arg); read(arg, *) chiId
    call get_command_argument(8, arg); read(arg, *) moleculeIntCode
    call get_command_argument(9, outDir)
    maxLevels = 1000000
    shapeId = 0
    if (command_argument_count() >= 10) then
        call get_command_argument(10, arg); read(arg, *) maxLevels
    end if
    if (command_argument_count() >= 11) then
        call get_command_argument(11, arg); read(arg, *) shapeId
    end if

    call readAtmosphere(trim(atmFile))
    call setShapeFunction()
    if (shapeId == 1) shapeFuncPtr => lorentzShape
    call setChiFactorFunction()
    select case (chiId)
    case (1)
        chiFactorFuncPtr => tonkov
    case (2)
        chiFactorFuncPtr => pollack
    case (3)
        chiFactorFuncPtr => perrin
    end select
    call readTIPS_local()

    open(databaseFileUnit, access='DIRECT', form='UNFORMATTED', recl=36, file=trim(dbFile))
    ! starting line as in app/main.f90:206-240
    startIdx = 1
    read(databaseFileUnit, rec=startIdx) startLineWV
    if (startWV > cutOff) then
        extStartWV = startWV - cutOff
        do while (startLineWV <= extStartWV)
            startIdx = startIdx + 1
            read(databaseFileUnit, rec=startIdx) startLineWV
        end do
    end if

    call cpu_time(t0)
    do lvl = 1, min(levels, maxLevels)
        pressure = pressureArray(lvl)
        temperature = temperatureArray(lvl)
        density = densityArray(lvl)
        pSelf = density * 10. * BOLsi * temperature / standardAtmosphericPressure
        pForeign = pressure - pSelf
        levelLabel = '___'
        if (lvl < 10) then
            write(levelLabel(1:1), '(I1)') lvl
        else if (lvl < 100) then
            write(levelLabel(1:2), '(I2)') lvl
        else
            write(levelLabel(1:3), '(I3)') lvl
        end if
        open(outUnit, access='DIRECT', form='UNFORMATTED', recl=NT*4, &
             file=trim(outDir)//'/'//levelLabel//'.ptbin')
        startDeltaWV = startWV
        endDeltaWV = startDeltaWV + deltaWV
        first = .true.
        do while (startDeltaWV < endWV)
            recNum = (startDeltaWV + 1.0) / 10.0
            extStartDeltaWV = startDeltaWV - cutOff
            extEndDeltaWV = startDeltaWV + deltaWV + cutOff
            if (first) then
                lineIdx = startIdx
                first = .false.
            end if
            capWV = extStartDeltaWV + deltaWV
            if (capWV <= startLineWV) capWV = startLineWV
            call resetAbsorptionGridValues()
            call lblScheme(lineIdx, capWV)
            call cascadeInterpolation()
            if (trim(target) == 'VAC') then
                write(outUnit, rec=recNum) density * RK
            else
                write(outUnit, rec=recNum) RK
            end if
            startDeltaWV = startDeltaWV + deltaWV
            endDeltaWV = startDeltaWV + deltaWV
        end do
        close(outUnit)
    end do
    call cpu_time(t1)
    write(*, '(A,F12.4)') 'LBL_SECONDS ', t1 - t0

contains

    subroutine readAtmosphere(fn)
        ! same layout as src/Atmsosphere.f90:29-60
        character(len=*), intent(in) :: fn
        integer :: k
        open(atmProfileUnit, file=fn, status='old')
        read(atmProfileUnit, '(A20)') atmTitle
        read(atmProfileUnit, *) levels
        allocate(heightArray(levels), pressureArray(levels), temperatureArray(levels), densityArray(levels))
        do k = 1, levels
            read(atmProfileUnit, *) heightArray(k), pressureArray(k), temperatureArray(k), densityArray(k)
        end do
        close(atmProfileUnit)
    end subroutine readAtmosphere

    subroutine readTIPS_local()
        integer :: nIso, nTemp, a, b
        open(unit=5467, file='data/TIPS/TIPS.dat', status='old', action='read')
        read(5467, *) nIso, nTemp
        allocate(TIPS(nIso, nTemp))
        do a = 1, nIso
            read(5467, *) (TIPS(a, b), b = 1, nTemp)
        end do
        close(5467)
    end subroutine readTIPS_local

end program ref_driver
