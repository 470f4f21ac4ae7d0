! marfa_par.f90 -- drop-in for the reference's .par preprocessing program
! (app/processParFile.f90): the same three arguments (molecule, database slug, .par file),
! the same outputs (data/databases/<slug>.<code> with 36-byte direct-access line records,
! <slug>.<code>.dat with the same values in list-directed text) and the same stop codes
! (1 wrong argument count, 2 unsupported molecule, 3 unreadable field).  The conversion
! is the library's marfa_par_convert (include/marfa.h); this program only handles the
! command line and writes the text copy from the records.
program marfa_par
    use, intrinsic :: iso_c_binding
    implicit none
    interface
        integer(c_int) function marfa_species_code(species, code) bind(C, name="marfa_species_code")
            import :: c_int, c_char
            character(kind=c_char), intent(in) :: species(*)
            character(kind=c_char), intent(out) :: code(*)
        end function marfa_species_code

        integer(c_int) function marfa_par_convert(species, par_path, db_path, nrec) &
                bind(C, name="marfa_par_convert")
            import :: c_int, c_char, c_int64_t
            character(kind=c_char), intent(in) :: species(*), par_path(*), db_path(*)
            integer(c_int64_t), intent(out) :: nrec
        end function marfa_par_convert
    end interface

    integer, parameter :: DP = selected_real_kind(15, 307)
    character(len=10) :: inputMolecule
    character(len=20) :: databaseSlug
    character(len=100) :: rawSpectralFile
    character(len=256) :: dbFile, datFile
    character(kind=c_char) :: code(3)
    character(len=2) :: nameSuffix
    integer(c_int) :: rc
    integer(c_int64_t) :: nrec, r
    integer :: argc, ios
    integer, parameter :: dbUnit = 341, datUnit = 744
    real(kind=DP) :: lineWV
    real :: S, gammaForeign, gammaSelf, lineLowerState, foreignTempCoeff, deltaForeign
    integer :: jointMolIso

    argc = command_argument_count()
    if (argc /= 3) then
        print *, 'Incorrect number of arguments.'
        print *, 'Expected 3 arguments, but received ', argc
        print *, 'Usage: program_name inputMolecule databaseSlug rawSpectralFile'
        stop 1
    end if
    call get_command_argument(1, inputMolecule)
    call get_command_argument(2, databaseSlug)
    call get_command_argument(3, rawSpectralFile)

    rc = marfa_species_code(trim(inputMolecule)//c_null_char, code)
    if (rc /= 0) then
        print *, 'ERROR: unsupported molecule: ', trim(adjustl(inputMolecule))
        stop 2
    end if
    nameSuffix = code(1)//code(2)

    dbFile = 'data/databases/'//trim(adjustl(databaseSlug))//'.'//nameSuffix
    datFile = trim(adjustl(databaseSlug))//'.'//nameSuffix//'.dat'
    rc = marfa_par_convert(trim(inputMolecule)//c_null_char, trim(rawSpectralFile)//c_null_char, &
                           trim(dbFile)//c_null_char, nrec)
    if (rc == 3) then
        print *, 'ERROR: when reading file:', trim(rawSpectralFile)
        stop 3
    else if (rc /= 0) then
        print *, 'ERROR: cannot convert ', trim(rawSpectralFile), ' into ', trim(dbFile), ' (code ', rc, ')'
        error stop
    end if

    ! text copy of the records (records left unwritten by skipped lines hold zeros)
    open(dbUnit, access='DIRECT', form='UNFORMATTED', recl=36, file=trim(dbFile), status='old', action='read')
    open(datUnit, file=trim(datFile), status='replace', action='write')
    do r = 1, nrec
        read(dbUnit, rec=r, iostat=ios) lineWV, S, gammaForeign, gammaSelf, lineLowerState, &
                                         foreignTempCoeff, jointMolIso, deltaForeign
        if (ios /= 0) exit
        if (lineWV == 0.0_DP .and. jointMolIso == 0) cycle
        write(datUnit, *) lineWV, S, gammaForeign, gammaSelf, lineLowerState, foreignTempCoeff, &
                          jointMolIso, deltaForeign
    end do
    close(dbUnit)
    close(datUnit)
    print *, 'The file preprocession is finished !'
    print *, 'number of lines: ', nrec
end program marfa_par
