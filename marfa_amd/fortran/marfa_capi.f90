! marfa_capi.f90 -- ISO_C_BINDING interface to the MI355X LBL core (include/marfa.h).
!
! One Fortran interface per C entry point; argument meaning and error codes as documented
! in include/marfa.h.  The derived type marfa_job_t mirrors the C struct marfa_job.
module marfa_capi
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: MARFA_NT = 20481
    integer(c_int), parameter :: MARFA_FP64 = 0, MARFA_FP32K = 1
    integer(c_int), parameter :: MARFA_ACS = 0, MARFA_VAC = 1
    integer(c_int), parameter :: MARFA_SHAPE_VOIGT = 0, MARFA_SHAPE_LORENTZ = 1
    integer(c_int), parameter :: MARFA_CHI_NONE = 0, MARFA_CHI_TONKOV = 1, MARFA_CHI_POLLACK = 2, &
                                 MARFA_CHI_PERRIN = 3

    type, bind(C) :: marfa_job_t
        real(c_double) :: start_wv
        real(c_double) :: end_wv
        real(c_float) :: cutoff
        integer(c_int32_t) :: shape_id
        integer(c_int32_t) :: chi_id
        integer(c_int32_t) :: mol_code
        integer(c_int32_t) :: target
        integer(c_int32_t) :: precision
    end type marfa_job_t

    interface
        integer(c_int) function marfa_ctx_create(device, ctx) bind(C, name="marfa_ctx_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
        end function marfa_ctx_create

        integer(c_int) function marfa_ctx_destroy(ctx) bind(C, name="marfa_ctx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function marfa_ctx_destroy

        type(c_ptr) function marfa_last_error(ctx) bind(C, name="marfa_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function marfa_last_error

        integer(c_int) function marfa_set_tips(ctx, niso, ntemp, tips) bind(C, name="marfa_set_tips")
            import :: c_int, c_int32_t, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: niso, ntemp
            real(c_float), intent(in) :: tips(*)
        end function marfa_set_tips

        integer(c_int) function marfa_set_lines(ctx, n, nu, S, gair, gself, elow, nT, delta, molIso) &
                bind(C, name="marfa_set_lines")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: nu(*)
            real(c_float), intent(in) :: S(*), gair(*), gself(*), elow(*), nT(*), delta(*)
            integer(c_int32_t), intent(in) :: molIso(*)
        end function marfa_set_lines

        integer(c_int) function marfa_set_levels(ctx, nlev, P, T, rho) bind(C, name="marfa_set_levels")
            import :: c_int, c_int32_t, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nlev
            real(c_float), intent(in) :: P(*), T(*), rho(*)
        end function marfa_set_levels

        integer(c_int) function marfa_windows(ctx, job, max_rec, nrec, lo, hi, recnum, origin) &
                bind(C, name="marfa_windows")
            import :: c_int, c_int32_t, c_ptr, marfa_job_t
            type(c_ptr), value :: ctx
            type(marfa_job_t), intent(in) :: job
            integer(c_int32_t), value :: max_rec
            integer(c_int32_t), intent(out) :: nrec
            type(c_ptr), value :: lo, hi, recnum, origin
        end function marfa_windows

        integer(c_int) function marfa_lbl_scheme(ctx, line_idx, cap_wv, start_dwv, ext_end_dwv, cutoff, &
                P, T, pSelf, shape_id, chi_id, mol_code, precision, rk) bind(C, name="marfa_lbl_scheme")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(inout) :: line_idx
            real(c_double), value :: cap_wv, start_dwv, ext_end_dwv
            real(c_float), value :: cutoff, P, T, pSelf
            integer(c_int32_t), value :: shape_id, chi_id, mol_code, precision
            real(c_float), intent(out) :: rk(*)
        end function marfa_lbl_scheme

        integer(c_int) function marfa_pt_tables(ctx, job, out_dev, out_host, stream, async) &
                bind(C, name="marfa_pt_tables")
            import :: c_int, c_int32_t, c_ptr, c_float, marfa_job_t
            type(c_ptr), value :: ctx
            type(marfa_job_t), intent(in) :: job
            type(c_ptr), value :: out_dev
            real(c_float), intent(out) :: out_host(*)
            type(c_ptr), value :: stream
            integer(c_int32_t), value :: async
        end function marfa_pt_tables
    end interface

contains

    ! message of the last failed call, as a Fortran string
    function marfa_error_text(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=256) :: msg
        character(kind=c_char), pointer :: p(:)
        type(c_ptr) :: cp
        integer :: i
        msg = ''
        cp = marfa_last_error(ctx)
        if (.not. c_associated(cp)) return
        call c_f_pointer(cp, p, [256])
        do i = 1, 256
            if (p(i) == c_null_char) exit
            msg(i:i) = p(i)
        end do
    end function marfa_error_text

end module marfa_capi
