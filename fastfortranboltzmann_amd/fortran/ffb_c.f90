! ffb_c.f90 -- ISO_C_BINDING interface to libffb_hip.so (include/ffb.h) for
! Fortran hosts.  A host `use ffb_c` and calls the C ABI directly; every
! function returns a status (0 = ok) and ffb_check() turns a failure into
! `error stop <message>`, the failure mode of the reference
! (source/layouts/xyd/simulation.f90:59-60).
module ffb_c
    use iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_int, c_ptr, c_null_ptr, &
        c_char, c_associated, c_f_pointer, c_loc
    implicit none
    private

    public :: ffb_config_t, ffb_config_default, ffb_create, ffb_destroy, ffb_init, ffb_step, &
        ffb_sync, ffb_read_macros, ffb_read_populations, ffb_elapsed_ms, ffb_reset_timer, &
        ffb_moments, ffb_deviation_extent, ffb_default_scale_log2, ffb_check, ffb_error_message, &
        ffb_device_properties, ffb_memory_usage
    public :: FFB_SHEAR_WAVE, FFB_COUETTE_FLOW, FFB_POISEUILLE_FLOW, FFB_SLIDING_LID
    public :: FFB_PREC_FP32, FFB_PREC_FP64, FFB_PREC_FP16_DEV, FFB_PREC_BF16_DEV, FFB_PREC_FP32_DEV, &
        FFB_PREC_E3M12_DEV
    public :: FFB_ROUND_NEAREST, FFB_ROUND_STOCHASTIC, FFB_E_DENSITY

    integer(c_int32_t), parameter :: FFB_SHEAR_WAVE = 1, FFB_COUETTE_FLOW = 2, &
        FFB_POISEUILLE_FLOW = 3, FFB_SLIDING_LID = 4
    integer(c_int32_t), parameter :: FFB_PREC_FP32 = 0, FFB_PREC_FP64 = 1, &
        FFB_PREC_FP16_DEV = 2, FFB_PREC_BF16_DEV = 3, FFB_PREC_FP32_DEV = 4, FFB_PREC_E3M12_DEV = 5
    integer(c_int32_t), parameter :: FFB_ROUND_NEAREST = 0, FFB_ROUND_STOCHASTIC = 1
    integer(c_int), parameter :: FFB_E_DENSITY = 4

    ! ffb_config_t, field for field
    type, bind(c) :: ffb_config_t
        integer(c_int32_t) :: n_x
        integer(c_int32_t) :: n_y
        integer(c_int32_t) :: sim_mode
        integer(c_int32_t) :: precision
        real(c_double) :: omega
        real(c_double) :: rho_0
        real(c_double) :: u_max
        real(c_double) :: n_sin
        real(c_double) :: u_wall
        real(c_double) :: u_lid
        real(c_double) :: rho_in
        real(c_double) :: rho_out
        integer(c_int32_t) :: storage_scale_log2
        integer(c_int32_t) :: device
        real(c_double) :: body_force
        integer(c_int32_t) :: storage_rounding
        integer(c_int32_t) :: density_checks
    end type ffb_config_t

    interface
        subroutine ffb_config_default(cfg) bind(c, name="ffb_config_default")
            import :: ffb_config_t
            type(ffb_config_t), intent(out) :: cfg
        end subroutine ffb_config_default

        integer(c_int) function ffb_create(cfg, lattice) bind(c, name="ffb_create")
            import :: ffb_config_t, c_int, c_ptr
            type(ffb_config_t), intent(in) :: cfg
            type(c_ptr), intent(out) :: lattice
        end function ffb_create

        integer(c_int) function ffb_destroy(lattice) bind(c, name="ffb_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: lattice
        end function ffb_destroy

        integer(c_int) function ffb_init(lattice) bind(c, name="ffb_init")
            import :: c_int, c_ptr
            type(c_ptr), value :: lattice
        end function ffb_init

        integer(c_int) function ffb_step(lattice, n_steps, write_macro_last) bind(c, name="ffb_step")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: lattice
            integer(c_int32_t), value :: n_steps
            integer(c_int32_t), value :: write_macro_last
        end function ffb_step

        integer(c_int) function ffb_sync(lattice) bind(c, name="ffb_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: lattice
        end function ffb_sync

        integer(c_int) function ffb_read_macros(lattice, rho, u_x, u_y) bind(c, name="ffb_read_macros")
            import :: c_int, c_ptr
            type(c_ptr), value :: lattice
            type(c_ptr), value :: rho
            type(c_ptr), value :: u_x
            type(c_ptr), value :: u_y
        end function ffb_read_macros

        integer(c_int) function ffb_read_populations(lattice, f_halo) bind(c, name="ffb_read_populations")
            import :: c_int, c_ptr
            type(c_ptr), value :: lattice
            type(c_ptr), value :: f_halo
        end function ffb_read_populations

        integer(c_int) function ffb_elapsed_ms(lattice, ms) bind(c, name="ffb_elapsed_ms")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lattice
            real(c_double), intent(out) :: ms
        end function ffb_elapsed_ms

        integer(c_int) function ffb_reset_timer(lattice) bind(c, name="ffb_reset_timer")
            import :: c_int, c_ptr
            type(c_ptr), value :: lattice
        end function ffb_reset_timer

        integer(c_int) function ffb_moments(lattice, mass, mom_x, mom_y) bind(c, name="ffb_moments")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lattice
            real(c_double), intent(out) :: mass, mom_x, mom_y
        end function ffb_moments

        ! largest |f - w| per direction class (rest, axis, diagonal) and the
        ! storage form's limit for it
        integer(c_int) function ffb_deviation_extent(lattice, max_dev, limit) &
                bind(c, name="ffb_deviation_extent")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lattice
            real(c_double), intent(out) :: max_dev(3), limit(3)
        end function ffb_deviation_extent

        ! default storage_scale_log2 of a form for flows of the given largest speed
        integer(c_int32_t) function ffb_default_scale_log2(precision, speed) &
                bind(c, name="ffb_default_scale_log2")
            import :: c_int32_t, c_double
            integer(c_int32_t), value :: precision
            real(c_double), value :: speed
        end function ffb_default_scale_log2

        ! name, compute units and memory of a HIP device (hardware report)
        integer(c_int) function ffb_device_properties(device, name, name_len, compute_units, memory_bytes) &
                bind(c, name="ffb_device_properties")
            import :: c_int, c_int32_t, c_int64_t, c_char
            integer(c_int32_t), value :: device
            character(kind=c_char), intent(out) :: name(*)
            integer(c_int32_t), value :: name_len
            integer(c_int32_t), intent(out) :: compute_units
            integer(c_int64_t), intent(out) :: memory_bytes
        end function ffb_device_properties

        ! device bytes of the handle: population sets, macro fields, the rest
        integer(c_int) function ffb_memory_usage(lattice, population_bytes, macro_bytes, other_bytes) &
                bind(c, name="ffb_memory_usage")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: lattice
            integer(c_int64_t), intent(out) :: population_bytes, macro_bytes, other_bytes
        end function ffb_memory_usage

        type(c_ptr) function c_ffb_last_error() bind(c, name="ffb_last_error")
            import :: c_ptr
        end function c_ffb_last_error

        integer(c_int) function c_strlen(s) bind(c, name="strlen")
            import :: c_int, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

contains

    ! Message of the last failing call (ffb_last_error) as a Fortran string.
    function ffb_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i

        p = c_ffb_last_error()
        if (.not. c_associated(p)) then
            msg = ""
            return
        end if
        n = c_strlen(p)
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function ffb_error_message

    ! error stop with the library's message on a non-zero status.
    subroutine ffb_check(status, what)
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: what

        if (status /= 0) then
            error stop "error: " // what // ": " // ffb_error_message()
        end if
    end subroutine ffb_check

end module ffb_c
