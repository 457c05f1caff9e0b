! ref_driver.f90 -- single-domain driver around the reference's own hot-path
! modules, used only to write golden vectors for the oracle (test
! infrastructure; never shipped, never run on a GPU).
!
! Compiled by oracle/build_ref.sh together with the UNMODIFIED reference files
!   source/settings.f90, source/layouts/xyd/modes/*.f90,
!   source/layouts/xyd/initialization.f90, source/layouts/xyd/simulation.f90
! read in place from the reference tree.  The reference's own driver
! (source/layouts/xyd/main.f90) and its exchange module
! (source/layouts/xyd/exchange.f90) need coarrays, which flang cannot lower,
! so this file restates the one-domain parts of them:
!   * the time loop, ping-pong and Poiseuille pre-pass order of main.f90:194-262
!   * the Poiseuille strip seeds of main.f90:139-155
!   * exchange_halos() with every neighbour equal to the own image
!     (exchange.f90:213-291, self-copy branches) and the plan of
!     build_exchange_plan() (exchange.f90:67-114) for one domain
!   * exchange_poiseuille_macro_halos() self-copy (exchange.f90:416-431)
! The domain_t and halo_buffers_t types come from the reference's own module
! files (domain.f90, exchange.f90 are parsed by flang for their .mod files).
!
! Usage: ref_driver <output.bin | ->
! Prints the wall time of the step loop and MLUPS = N_X*N_Y*N_STEPS/s/1e6
! (main.f90:287-326).  With "-" no state is written.
! Output (stream, native endian): f(0:N_X+1,0:N_Y+1,9) of the current buffer
! after N_STEPS steps, then rho, u_x, u_y (N_X,N_Y) of the last step, then
! recv_macro_left, recv_macro_right (N_Y,3) (zeros unless Poiseuille).
program ref_driver
    use iso_fortran_env, only: int32, int64, real64
    use domain, only: domain_t
    use exchange, only: halo_buffers_t
    use settings, only: N_X, N_Y, N_STEPS, N_DIRS, FP, SIM_MODE, &
        SIM_SHEAR_WAVE, SIM_COUETTE_FLOW, SIM_POISEUILLE_FLOW, SIM_SLIDING_LID, &
        RHO_0, U_MAX, N_SIN, RHO_IN, RHO_OUT
    use initialization, only: apply_condition_shear_wave_local, &
        apply_condition_couette_flow_local, apply_condition_poiseuille_flow_local, &
        apply_condition_sliding_lid_local
    use poiseuille_flow, only: prepare_poiseuille_flow_halos_PF, update_poiseuille_flow_macro_strips_PF
    use simulation, only: execute_local_sim_step
    implicit none

    type(domain_t) :: dom
    type(halo_buffers_t) :: hb
    real(FP), allocatable, target :: f_a(:,:,:), f_b(:,:,:)
    real(FP), allocatable :: rho(:,:), u_x(:,:), u_y(:,:)
    real(FP), allocatable :: win_left(:,:), win_right(:,:)
    logical :: read_from_a, write_macro
    integer(int32) :: step, unit_out
    integer(int64) :: clock_start, clock_end, clock_rate
    real(real64) :: elapsed_seconds
    character(len=4096) :: out_path

    call get_command_argument(1, out_path)

    dom%n_images = 1
    dom%n_images_x = 1
    dom%n_images_y = 1
    dom%image_id = 1
    dom%n_x = N_X
    dom%n_y = N_Y
    dom%y_global_start = 1
    dom%at_left_boundary = .true.
    dom%at_right_boundary = .true.
    dom%at_bottom_boundary = .true.
    dom%at_top_boundary = .true.

    allocate(f_a(0:N_X+1, 0:N_Y+1, N_DIRS), f_b(0:N_X+1, 0:N_Y+1, N_DIRS))
    allocate(rho(N_X, N_Y), u_x(N_X, N_Y), u_y(N_X, N_Y))
    allocate(hb%recv_macro_left(N_Y, 3), hb%recv_macro_right(N_Y, 3))
    allocate(win_left(N_Y, 3), win_right(N_Y, 3))
    f_b = 0.0_FP
    hb%recv_macro_left = 0.0_FP
    hb%recv_macro_right = 0.0_FP

    select case (SIM_MODE)
    case (SIM_SHEAR_WAVE)
        call apply_condition_shear_wave_local(RHO_0, U_MAX, N_SIN, N_X, N_Y, 1_int32, f_a, rho, u_x, u_y)
    case (SIM_COUETTE_FLOW)
        call apply_condition_couette_flow_local(RHO_0, N_X, N_Y, f_a, rho, u_x, u_y)
    case (SIM_POISEUILLE_FLOW)
        call apply_condition_poiseuille_flow_local(RHO_0, N_X, N_Y, f_a, rho, u_x, u_y)
    case (SIM_SLIDING_LID)
        call apply_condition_sliding_lid_local(RHO_0, N_X, N_Y, f_a, rho, u_x, u_y)
    end select

    if (SIM_MODE == SIM_POISEUILLE_FLOW) then
        win_left(:, 1) = RHO_IN
        win_left(:, 2:3) = 0.0_FP
        win_right(:, 1) = RHO_OUT
        win_right(:, 2:3) = 0.0_FP
        hb%recv_macro_left(:, 1) = RHO_IN
        hb%recv_macro_left(:, 2:3) = 0.0_FP
        hb%recv_macro_right(:, 1) = RHO_OUT
        hb%recv_macro_right(:, 2:3) = 0.0_FP
    end if

    read_from_a = .true.
    call system_clock(clock_start, clock_rate)
    do step = 1, N_STEPS
        write_macro = (step == N_STEPS)
        if (read_from_a) then
            call one_step(f_a, f_b)
        else
            call one_step(f_b, f_a)
        end if
        read_from_a = .not. read_from_a
    end do
    call system_clock(clock_end)
    elapsed_seconds = real(clock_end - clock_start, real64) / real(clock_rate, real64)
    write(*, "(a, i0, a, i0, a, es14.6, a, f12.3)") "ref_driver: steps=", N_STEPS, " cells=", &
        int(N_X, int64) * int(N_Y, int64), " seconds=", elapsed_seconds, " MLUPS=", &
        real(N_X, real64) * real(N_Y, real64) * real(N_STEPS, real64) / elapsed_seconds / 1.0e6_real64

    if (trim(out_path) == "-") stop

    open(newunit=unit_out, file=trim(out_path), access="stream", form="unformatted", status="replace")
    if (read_from_a) then
        write(unit_out) f_a
    else
        write(unit_out) f_b
    end if
    write(unit_out) rho, u_x, u_y
    write(unit_out) hb%recv_macro_left, hb%recv_macro_right
    close(unit_out)

contains

    subroutine one_step(f, f_next)
        real(FP), intent(inout) :: f(0:N_X+1, 0:N_Y+1, N_DIRS)
        real(FP), intent(inout) :: f_next(0:N_X+1, 0:N_Y+1, N_DIRS)

        call exchange_one_domain(f)

        if (SIM_MODE == SIM_POISEUILLE_FLOW) then
            call prepare_poiseuille_flow_halos_PF(N_X, N_Y, .true., .true., .true., .true., &
                RHO_IN, RHO_OUT, f, hb%recv_macro_left, hb%recv_macro_right)
            call update_poiseuille_flow_macro_strips_PF(N_X, N_Y, .true., .true., f, win_left, win_right)
            hb%recv_macro_left = win_right
            hb%recv_macro_right = win_left
        end if

        call execute_local_sim_step(dom, hb, N_X, N_Y, write_macro, f, f_next, rho, u_x, u_y)
    end subroutine one_step

    subroutine exchange_one_domain(f)
        real(FP), intent(inout) :: f(0:N_X+1, 0:N_Y+1, N_DIRS)
        logical :: plan_x, plan_y

        plan_x = SIM_MODE == SIM_SHEAR_WAVE .or. SIM_MODE == SIM_COUETTE_FLOW .or. &
            SIM_MODE == SIM_POISEUILLE_FLOW
        plan_y = SIM_MODE == SIM_SHEAR_WAVE

        if (plan_x) then
            f(0, 1:N_Y, 2) = f(N_X, 1:N_Y, 2)
            f(0, 1:N_Y, 6) = f(N_X, 1:N_Y, 6)
            f(0, 1:N_Y, 9) = f(N_X, 1:N_Y, 9)
            f(N_X+1, 1:N_Y, 4) = f(1, 1:N_Y, 4)
            f(N_X+1, 1:N_Y, 7) = f(1, 1:N_Y, 7)
            f(N_X+1, 1:N_Y, 8) = f(1, 1:N_Y, 8)
        end if
        if (plan_y) then
            f(0:N_X+1, 0, 3) = f(0:N_X+1, N_Y, 3)
            f(0:N_X+1, 0, 6) = f(0:N_X+1, N_Y, 6)
            f(0:N_X+1, 0, 7) = f(0:N_X+1, N_Y, 7)
            f(0:N_X+1, N_Y+1, 5) = f(0:N_X+1, 1, 5)
            f(0:N_X+1, N_Y+1, 8) = f(0:N_X+1, 1, 8)
            f(0:N_X+1, N_Y+1, 9) = f(0:N_X+1, 1, 9)
        end if
    end subroutine exchange_one_domain

end program ref_driver
