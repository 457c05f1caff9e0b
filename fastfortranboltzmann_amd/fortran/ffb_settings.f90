! ffb_settings.f90 -- run settings of the Fortran host.
!
! Same surface as the reference's compile-time settings
! (source/settings.f90:21-93): the FFB_* preprocessor definitions
! (FFB_USE_CMAKE_SETTINGS with FFB_N_X, FFB_N_Y, FFB_N_STEPS, FFB_SIM_MODE,
! FFB_RHO_0, FFB_OMEGA, FFB_U_MAX, FFB_N_SIN, FFB_U_WALL, FFB_U_LID, FFB_RHO_IN,
! FFB_RHO_OUT, FFB_EXPORT_*, FFB_PROGRESS_INTERVAL, ...; FFB_FP64 for the
! precision) set the defaults, with the reference's values when absent.
! Because the solver is a library, one binary serves every configuration:
! an environment variable of the same name (e.g. FFB_N_X=4096) overrides the
! compiled value at run time.  FFB_PRECISION=fp32|fp64|fp16dev|bf16dev|fp32dev|e3m12dev and
! FFB_STORAGE_SCALE_LOG2 (and FFB_STORAGE_ROUNDING=stochastic|nearest for the
! 16-bit forms) select the storage forms the reference lacks; without
! FFB_STORAGE_SCALE_LOG2 the scale is the library's default for the form and
! the mode's speed (ffb_default_scale_log2).
module ffb_settings
    use iso_fortran_env, only: int32, int64, real32, real64
    implicit none
    private

    ! kind of the compile-time settings literals (1.1_FP, ...), as the
    ! reference's FP (settings.f90:7-13): real32 unless FFB_FP64, so the values
    ! a settings build passes are the reference's constants exactly
#ifdef FFB_FP64
    integer, parameter :: FP = real64
#else
    integer, parameter :: FP = real32
#endif

    public :: run_settings_t, load_settings, mode_name, reynolds_number
    ! storage_scale_log2 not given: the library's default for the form and speed
    integer(int32), parameter, public :: STORAGE_SCALE_DEFAULT = -huge(1_int32)

    integer(int32), parameter, public :: SIM_SHEAR_WAVE = 1, SIM_COUETTE_FLOW = 2, &
        SIM_POISEUILLE_FLOW = 3, SIM_SLIDING_LID = 4

    type :: run_settings_t
        integer(int32) :: n_x, n_y, n_steps, sim_mode
        real(real64) :: rho_0, omega, u_max, n_sin, u_wall, u_lid, rho_in, rho_out
        logical :: export_macros, export_endpoint_states, interactive_progress
        integer(int32) :: export_interval, progress_interval
        character(len=:), allocatable :: export_num
        character(len=:), allocatable :: precision
        integer(int32) :: storage_scale_log2
        character(len=:), allocatable :: storage_rounding
        logical :: density_checks
    end type run_settings_t

contains

    subroutine load_settings(s)
        type(run_settings_t), intent(out) :: s

#ifdef FFB_USE_CMAKE_SETTINGS
        s%n_x = FFB_N_X
        s%n_y = FFB_N_Y
        s%n_steps = FFB_N_STEPS
        s%sim_mode = FFB_SIM_MODE
        s%rho_0 = real(FFB_RHO_0, real64)
        s%omega = real(FFB_OMEGA, real64)
        s%u_max = real(FFB_U_MAX, real64)
        s%n_sin = real(FFB_N_SIN, real64)
        s%u_wall = real(FFB_U_WALL, real64)
        s%u_lid = real(FFB_U_LID, real64)
        s%rho_in = real(FFB_RHO_IN, real64)
        s%rho_out = real(FFB_RHO_OUT, real64)
        s%export_macros = FFB_EXPORT_MACROS
        s%export_endpoint_states = FFB_EXPORT_ENDPOINT_STATES
        s%export_interval = FFB_EXPORT_INTERVAL
        s%interactive_progress = FFB_INTERACTIVE_PROGRESS
        s%progress_interval = FFB_PROGRESS_INTERVAL
#else
        ! reference defaults, source/settings.f90:27-93
        s%n_x = 2520
        s%n_y = 2520
        s%n_steps = 1000
        s%sim_mode = SIM_SLIDING_LID
        s%rho_0 = 1.0_real64
        s%omega = 1.1_real64
        s%u_max = 0.1_real64
        s%n_sin = 2.0_real64
        s%u_wall = 0.1_real64
        s%u_lid = 0.1_real64
        s%rho_in = 1.00005_real64
        s%rho_out = 0.99995_real64
        s%export_macros = .false.
        s%export_endpoint_states = .true.
        s%export_interval = 10000
        s%interactive_progress = .false.
        s%progress_interval = 1000
#endif
#ifdef FFB_EXPORT_NUM
        s%export_num = FFB_EXPORT_NUM
#else
        s%export_num = "run_000"
#endif
#ifdef FFB_FP64
        s%precision = "fp64"
#else
        s%precision = "fp32"
#endif
        s%storage_scale_log2 = STORAGE_SCALE_DEFAULT
        s%storage_rounding = "stochastic"
        ! FFB_DENSITY_CHECKS: stop at a density rho <= 0 (simulation.f90:229-233)
#ifdef FFB_DENSITY_CHECKS
        s%density_checks = .true.
#else
        s%density_checks = .false.
#endif

        call env_int("FFB_N_X", s%n_x)
        call env_int("FFB_N_Y", s%n_y)
        call env_int("FFB_N_STEPS", s%n_steps)
        call env_int("FFB_SIM_MODE", s%sim_mode)
        call env_real("FFB_RHO_0", s%rho_0)
        call env_real("FFB_OMEGA", s%omega)
        call env_real("FFB_U_MAX", s%u_max)
        call env_real("FFB_N_SIN", s%n_sin)
        call env_real("FFB_U_WALL", s%u_wall)
        call env_real("FFB_U_LID", s%u_lid)
        call env_real("FFB_RHO_IN", s%rho_in)
        call env_real("FFB_RHO_OUT", s%rho_out)
        call env_logical("FFB_EXPORT_MACROS", s%export_macros)
        call env_logical("FFB_EXPORT_ENDPOINT_STATES", s%export_endpoint_states)
        call env_int("FFB_EXPORT_INTERVAL", s%export_interval)
        call env_logical("FFB_INTERACTIVE_PROGRESS", s%interactive_progress)
        call env_int("FFB_PROGRESS_INTERVAL", s%progress_interval)
        call env_string("FFB_EXPORT_NUM", s%export_num)
        call env_string("FFB_PRECISION", s%precision)
        call env_int("FFB_STORAGE_SCALE_LOG2", s%storage_scale_log2)
        call env_string("FFB_STORAGE_ROUNDING", s%storage_rounding)
        call env_logical("FFB_DENSITY_CHECKS", s%density_checks)

        if (s%n_x < 1 .or. s%n_y < 1) error stop "error: N_X and N_Y must be positive"
        if (s%n_steps < 0) error stop "error: N_STEPS must not be negative"
        if (s%sim_mode < 1 .or. s%sim_mode > 4) error stop "error: unknown sim mode"
        if (s%progress_interval < 1) s%progress_interval = 1
    end subroutine load_settings

    ! name strings of sim_mode_to_string, source/settings.f90:130-151
    function mode_name(sim_mode) result(name)
        integer(int32), intent(in) :: sim_mode
        character(len=:), allocatable :: name

        select case (sim_mode)
        case (SIM_SHEAR_WAVE)
            name = "shear_wave"
        case (SIM_COUETTE_FLOW)
            name = "couette_flow"
        case (SIM_POISEUILLE_FLOW)
            name = "poiseuille_flow"
        case (SIM_SLIDING_LID)
            name = "sliding_lid"
        case default
            name = "unknown_sim_mode"
        end select
    end function mode_name

    ! Re = U * N_Y / nu with nu = (1/omega - 1/2)/3, source/settings.f90:154-176
    function reynolds_number(s) result(re)
        type(run_settings_t), intent(in) :: s
        real(real64) :: re, nu

        nu = (1.0_real64 / s%omega - 0.5_real64) / 3.0_real64
        select case (s%sim_mode)
        case (SIM_COUETTE_FLOW)
            re = s%u_wall * real(s%n_y, real64) / nu
        case (SIM_SLIDING_LID)
            re = s%u_lid * real(s%n_y, real64) / nu
        case default
            re = 0.0_real64
        end select
    end function reynolds_number

    subroutine env_int(name, v)
        character(len=*), intent(in) :: name
        integer(int32), intent(inout) :: v
        character(len=64) :: buf
        integer :: n, st, ios

        call get_environment_variable(name, buf, n, st)
        if (st == 0 .and. n > 0) then
            read(buf(1:n), *, iostat=ios) v
            if (ios /= 0) error stop "error: bad integer in environment variable " // name
        end if
    end subroutine env_int

    subroutine env_real(name, v)
        character(len=*), intent(in) :: name
        real(real64), intent(inout) :: v
        character(len=64) :: buf
        integer :: n, st, ios, k

        call get_environment_variable(name, buf, n, st)
        if (st == 0 .and. n > 0) then
            k = index(buf(1:n), "_")
            if (k > 0) n = k - 1   ! accept Fortran literals such as 1.1_FP
            read(buf(1:n), *, iostat=ios) v
            if (ios /= 0) error stop "error: bad real in environment variable " // name
        end if
    end subroutine env_real

    subroutine env_logical(name, v)
        character(len=*), intent(in) :: name
        logical, intent(inout) :: v
        character(len=16) :: buf
        integer :: n, st

        call get_environment_variable(name, buf, n, st)
        if (st == 0 .and. n > 0) then
            select case (buf(1:n))
            case (".true.", "true", "1", "T", "ON", "on")
                v = .true.
            case (".false.", "false", "0", "F", "OFF", "off")
                v = .false.
            case default
                error stop "error: bad logical in environment variable " // name
            end select
        end if
    end subroutine env_logical

    subroutine env_string(name, v)
        character(len=*), intent(in) :: name
        character(len=:), allocatable, intent(inout) :: v
        character(len=256) :: buf
        integer :: n, st

        call get_environment_variable(name, buf, n, st)
        if (st == 0 .and. n > 0) v = buf(1:n)
    end subroutine env_string

end module ffb_settings
