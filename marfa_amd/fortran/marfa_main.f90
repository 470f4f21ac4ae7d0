! marfa_main.f90 -- PT-table generator on the MI355X LBL core: drop-in for app/main.f90.
!
! Same command line, input files, output layout and stop codes as the reference program:
!   marfa_gpu Molecule StartWV EndWV DatabaseSlug CutOff TargetValue AtmProfileFile
!             [Shape [Chi [Precision]]]
! reads data/databases/<DatabaseSlug>.<code> (36-byte direct-access line records),
! data/Atmospheres/<AtmProfileFile> and data/TIPS/TIPS.dat relative to the working
! directory, and writes output/ptTables/<Molecule>_<StartWV>-<EndWV>_<timestamp>/<level>.ptbin
! (one 20,481-float record per 10 cm-1 subinterval, record number int((start+1)/10)),
! info.txt and output/ptTables/latest_run.txt.
! The three optional arguments choose what the reference fixes at compile time:
!   Shape     VOIGT (default) | LORENTZ      -- setShapeFunction, src/Shapes.f90:11-15
!   Chi       NONE (default) | TONKOV | POLLACK | PERRIN -- setChiFactorFunction,
!             src/ChiFactors.f90:11-16
!   Precision FP64 (default) | FP32K         -- arithmetic of the device path
! Instead of looping over levels and subintervals on the CPU (app/main.f90:81-180), the
! whole job runs on the GPU through marfa_pt_tables, a batch of levels at a time; the
! host writes the returned records.  Library errors end the program with the library's
! error code as stop code (9 = line-data error, as src/LBL.f90:50-53).
program marfa_gpu_main
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: int8, int64
    use marfa_capi
    implicit none

    integer, parameter :: dp = kind(1.d0)
    integer, parameter :: out_unit = 47, db_unit = 7777, atm_unit = 5555, tips_unit = 7788
    integer, parameter :: info_unit = 67, latest_unit = 87
    integer(int64), parameter :: host_budget = 2_int64 * 1024_int64**3  ! bytes of records per batch

    character(len=10) :: molecule, a_start, a_end, a_cut
    character(len=20) :: db_slug
    character(len=30) :: target_s, shape_s, chi_s, prec_s
    character(len=200) :: atm_file
    character(len=2) :: db_ext
    character(len=20) :: atm_title
    character(len=300) :: sub_dir, full_dir, db_path
    character(len=3) :: label
    character(len=20) :: stamp
    integer :: dtv(8), argc, st, ios, nlev, lev, l0, nb, i, ir, nrec_i, mol_code
    integer(c_int32_t) :: nrec, niso, ntemp
    integer(int64) :: nlines, fsize, per_level
    real(dp) :: start_wv, end_wv
    real :: cut_off
    real, allocatable :: hgt(:), p_arr(:), t_arr(:), rho_arr(:), tips(:, :), recs(:)
    real(c_double), allocatable :: nu(:)
    real(c_double), allocatable, target :: origin(:)
    real(c_float), allocatable :: s(:), gair(:), gself(:), elow(:), ntc(:), delta(:)
    integer(c_int32_t), allocatable :: moliso(:)
    integer(c_int32_t), allocatable, target :: recnum(:)
    integer(c_int64_t), allocatable, target :: lo(:), hi(:)
    type(c_ptr) :: ctx
    type(marfa_job_t), target :: job
    integer(int64) :: t0, t1, rate

    call system_clock(t0, rate)

    ! ---- command line (src/EnvironmentSetup.f90:34-95) ----
    argc = command_argument_count()
    if (argc < 7) then
        print *, 'Insufficient number of arguments.'
        print *, 'Expected at least 7 arguments, but received ', argc
        print *, 'Usage: marfa_gpu Molecule StartWV EndWV DatabaseSlug CutOff TargetValue ', &
                 'AtmProfileFile [Shape [Chi [Precision]]]'
        print *, 'TargetValue: ACS (cross-section, cm^2/molecule) or VAC (volume absorption, km-1)'
        print *, 'Shape: VOIGT | LORENTZ;  Chi: NONE | TONKOV | POLLACK | PERRIN;  Precision: FP64 | FP32K'
        stop 1
    end if
    call get_command_argument(1, molecule)
    call get_command_argument(2, a_start)
    read(a_start, *) start_wv
    call get_command_argument(3, a_end)
    read(a_end, *) end_wv
    call get_command_argument(4, db_slug)
    call get_command_argument(5, a_cut)
    read(a_cut, *) cut_off
    call get_command_argument(6, target_s)
    call get_command_argument(7, atm_file)
    shape_s = 'VOIGT'
    chi_s = 'NONE'
    prec_s = 'FP64'
    if (argc >= 8) call get_command_argument(8, shape_s)
    if (argc >= 9) call get_command_argument(9, chi_s)
    if (argc >= 10) call get_command_argument(10, prec_s)

    job%start_wv = start_wv
    job%end_wv = end_wv
    job%cutoff = cut_off
    select case (trim(target_s))
    case ('ACS')
        job%target = MARFA_ACS
    case ('VAC')
        job%target = MARFA_VAC
    case default
        print *, 'Error: Invalid targetValue "', trim(target_s), '". Must be either "ACS" ', &
                 '(Absorption cross-section) or "VAC" (Volume absorption coefficient).'
        stop 2
    end select
    select case (trim(shape_s))
    case ('VOIGT')
        job%shape_id = MARFA_SHAPE_VOIGT
    case ('LORENTZ')
        job%shape_id = MARFA_SHAPE_LORENTZ
    case default
        print *, 'Error: unknown shape "', trim(shape_s), '" (VOIGT or LORENTZ)'
        stop 2
    end select
    select case (trim(chi_s))
    case ('NONE')
        job%chi_id = MARFA_CHI_NONE
    case ('TONKOV')
        job%chi_id = MARFA_CHI_TONKOV
    case ('POLLACK')
        job%chi_id = MARFA_CHI_POLLACK
    case ('PERRIN')
        job%chi_id = MARFA_CHI_PERRIN
    case default
        print *, 'Error: unknown chi factor "', trim(chi_s), '" (NONE, TONKOV, POLLACK or PERRIN)'
        stop 2
    end select
    select case (trim(prec_s))
    case ('FP64')
        job%precision = MARFA_FP64
    case ('FP32K')
        job%precision = MARFA_FP32K
    case default
        print *, 'Error: unknown precision "', trim(prec_s), '" (FP64 or FP32K)'
        stop 2
    end select

    ! ---- atmosphere (src/Atmsosphere.f90:29-60) ----
    open(atm_unit, file='data/Atmospheres/'//trim(atm_file), status='old', iostat=ios)
    if (ios /= 0) then
        print *, 'Error: Unable to open "data/Atmospheres/', trim(atm_file), '".'
        stop 7
    end if
    read(atm_unit, '(A20)') atm_title
    read(atm_unit, *) nlev
    if (nlev > 200) then
        write(*, '(A, I3, A)') 'WARNING: input number of atmospheric levels is bigger than ', 200
        stop 6
    end if
    allocate(hgt(nlev), p_arr(nlev), t_arr(nlev), rho_arr(nlev))
    do lev = 1, nlev
        read(atm_unit, *, iostat=ios) hgt(lev), p_arr(lev), t_arr(lev), rho_arr(lev)
        if (ios /= 0) then
            print *, 'Error: Unable to read data from file "data/Atmospheres/', trim(atm_file), '".'
            print *, 'Check the data format in the atmospheric file'
            close(atm_unit)
            stop 7
        end if
    end do
    close(atm_unit)

    ! ---- TIPS (src/EnvironmentSetup.f90:169-184) ----
    open(tips_unit, file='data/TIPS/TIPS.dat', status='old', action='read')
    read(tips_unit, *) niso, ntemp
    allocate(tips(ntemp, niso))
    do i = 1, niso
        read(tips_unit, *) tips(:, i)
    end do
    close(tips_unit)

    ! ---- species code (src/EnvironmentSetup.f90:187-235) ----
    call species_code(molecule, mol_code, db_ext)
    if (mol_code == 0) then
        print *, 'ERROR: unsupported molecule: ', trim(adjustl(molecule))
        stop 8
    end if
    job%mol_code = mol_code

    ! ---- output directory and run info (src/EnvironmentSetup.f90:98-166) ----
    call date_and_time(values=dtv)
    write(stamp, '(I4, I2.2, I2.2, "_", I2.2, I2.2, I2.2, I2.2)') dtv(1), dtv(2), dtv(3), dtv(5), &
        dtv(6), dtv(7)
    write(a_start, '(I0)') int(start_wv)
    write(a_end, '(I0)') int(end_wv)
    sub_dir = trim(molecule)//'_'//trim(a_start)//'-'//trim(a_end)//'_'//trim(stamp)
    full_dir = 'output/ptTables/'//trim(sub_dir)
    call execute_command_line('mkdir -p "'//trim(full_dir)//'"', wait=.true., exitstat=st)
    if (st /= 0) then
        print *, 'Error: Failed to create directory ', trim(full_dir)
        stop 3
    end if
    print *, 'Directory for storing PT-tables is created: ', trim(full_dir)
    open(info_unit, file=trim(full_dir)//'/info.txt', status='replace', action='write', iostat=st)
    if (st /= 0) then
        print *, 'Error: Unable to create info file at ', trim(full_dir)//'/info.txt'
        stop 4
    end if
    write(info_unit, '(A)') 'Command-Line Arguments:'
    write(info_unit, '(A,A)') 'Input Molecule: ', trim(molecule)
    call get_command_argument(2, a_start)
    call get_command_argument(3, a_end)
    write(info_unit, '(A,A)') 'Start Wavenumber: ', trim(a_start)
    write(info_unit, '(A,A)') 'End Wavenumber: ', trim(a_end)
    write(info_unit, '(A,A)') 'Cut Off: ', trim(a_cut)
    write(info_unit, '(A,A)') 'Target Value: ', trim(target_s)
    write(info_unit, '(A,A)') 'Atmospheric Profile File: ', trim(atm_file)
    write(info_unit, '(A,A)') 'Shape: ', trim(shape_s)
    write(info_unit, '(A,A)') 'Chi Factor: ', trim(chi_s)
    write(info_unit, '(A,A)') 'Precision: ', trim(prec_s)
    close(info_unit)
    print *, 'Info file created at: ', trim(full_dir)//'/info.txt'
    open(latest_unit, file='output/ptTables/latest_run.txt', status='replace', action='write', &
         iostat=st)
    if (st /= 0) then
        print *, 'Error: Unable to create latest run file at output/ptTables/latest_run.txt'
        stop 5
    end if
    write(latest_unit, '(A)') trim(sub_dir)
    close(latest_unit)

    ! ---- line list: the whole database file, read as a byte stream ----
    db_path = 'data/databases/'//trim(adjustl(db_slug))//'.'//db_ext
    inquire(file=trim(db_path), size=fsize)
    if (fsize < 36) then
        print *, 'ERROR: when reading file with spectral data: ', trim(db_path)
        stop 9
    end if
    nlines = fsize / 36
    allocate(nu(nlines), s(nlines), gair(nlines), gself(nlines), elow(nlines), ntc(nlines), &
             delta(nlines), moliso(nlines))
    open(db_unit, file=trim(db_path), access='stream', form='unformatted', status='old', &
         action='read')
    read(db_unit, iostat=ios) (nu(i), s(i), gair(i), gself(i), elow(i), ntc(i), moliso(i), &
                               delta(i), i = 1, int(nlines))
    close(db_unit)
    if (ios /= 0) then
        print *, 'ERROR: when reading file with spectral data.'
        stop 9
    end if

    ! ---- device context and inputs ----
    call check(marfa_ctx_create(0_c_int, ctx), c_null_ptr)
    call check(marfa_set_tips(ctx, niso, ntemp, tips), ctx)
    call check(marfa_set_lines(ctx, nlines, nu, s, gair, gself, elow, ntc, delta, moliso), ctx)

    call check(marfa_windows(ctx, job, 0_c_int32_t, nrec, c_null_ptr, c_null_ptr, c_null_ptr, &
                             c_null_ptr), ctx)
    nrec_i = nrec
    allocate(lo(max(1, nrec_i)), hi(max(1, nrec_i)), recnum(max(1, nrec_i)), origin(max(1, nrec_i)))
    call check(marfa_windows(ctx, job, nrec, nrec, c_loc(lo), c_loc(hi), c_loc(recnum), &
                             c_loc(origin)), ctx)
    if (nrec_i > 0) then
        if (start_wv > cut_off .and. abs(start_wv - cut_off - nu(lo(1) + 1)) > 25.) then
            print *, 'ATTENTION: database file might be insufficient for input spectral interval:'
            print *, 'Your input left boundary - cutOff condition: ', start_wv - cut_off, ' cm-1'
            print *, 'Line-by-line scheme starts from: ', nu(lo(1) + 1), ' cm-1'
        end if
        do ir = 1, nrec_i
            if (recnum(ir) < 1) then
                print *, 'ERROR: subinterval starting at ', origin(ir), ' cm-1 maps to record ', &
                         recnum(ir)
                stop 10
            end if
        end do
    end if

    ! ---- all levels x subintervals on the GPU, a batch of levels at a time ----
    per_level = max(1_int64, int(nrec_i, int64) * MARFA_NT * 4_int64)
    nb = int(max(1_int64, min(int(nlev, int64), host_budget / per_level)))
    allocate(recs(int(nb, int64) * max(1, nrec_i) * MARFA_NT))
    do l0 = 1, nlev, nb
        lev = min(nlev, l0 + nb - 1)
        call check(marfa_set_levels(ctx, int(lev - l0 + 1, c_int32_t), p_arr(l0:lev), t_arr(l0:lev), &
                                    rho_arr(l0:lev)), ctx)
        if (nrec_i > 0) call check(marfa_pt_tables(ctx, job, c_null_ptr, recs, c_null_ptr, 0_c_int32_t), &
                                   ctx)
        do i = l0, lev
            label = '___'
            if (i < 10) then
                write(label(1:1), '(I1)') i
            else if (i < 100) then
                write(label(1:2), '(I2)') i
            else
                write(label(1:3), '(I3)') i
            end if
            open(out_unit, access='DIRECT', form='UNFORMATTED', recl=MARFA_NT * 4, &
                 file=trim(full_dir)//'/'//label//'.ptbin')
            do ir = 1, nrec_i
                write(out_unit, rec=recnum(ir)) &
                    recs((int(i - l0, int64) * nrec_i + ir - 1) * MARFA_NT + 1: &
                         (int(i - l0, int64) * nrec_i + ir) * MARFA_NT)
            end do
            close(out_unit)
            print *, i, ' of ', nlev, ' atmospheric levels is being processed ...'
        end do
    end do
    call check(marfa_ctx_destroy(ctx), c_null_ptr)

    print *, ' *** Congratulations! PT-tables have been calculated! ***'
    call system_clock(t1)
    print *, 'Took: ', real(t1 - t0, dp) / real(rate, dp), ' seconds'

contains

    subroutine check(code, c)
        integer(c_int), intent(in) :: code
        type(c_ptr), intent(in) :: c
        if (code == 0) return
        if (c_associated(c)) then
            print *, 'ERROR (marfa ', code, '): ', trim(marfa_error_text(c))
        else
            print *, 'ERROR (marfa ', code, ')'
        end if
        error stop code
    end subroutine check

    pure subroutine species_code(species, code, ext)
        character(len=*), intent(in) :: species
        integer, intent(out) :: code
        character(len=2), intent(out) :: ext
        character(len=4), parameter :: names(12) = [character(len=4) :: 'H2O', 'CO2', 'O3', 'N2O', &
            'CO', 'CH4', 'O2', 'NO', 'SO2', 'NO2', 'NH3', 'HNO3']
        integer :: k
        code = 0
        ext = '00'
        do k = 1, 12
            if (trim(adjustl(species)) == trim(names(k))) then
                code = k
                write(ext, '(I2.2)') k
                return
            end if
        end do
    end subroutine species_code

end program marfa_gpu_main
