! ffb_host.f90 -- Fortran host driving the MI355X solver through the C ABI.
!
! Counterpart of the reference's driver (source/layouts/xyd/main.f90) for one
! device-resident domain: the same settings surface (ffb_settings), the same
! stdout blocks that scripts/benchmark_run.py parses (reporting.f90:22-234:
! "[hh:mm:ss] launched", the execution-time spread table, "step time = X ms",
! "MLUPS = N") and the same export files (export.f90:92-524:
! output/<EXPORT_NUM>/config.json and {density,velocity_x,velocity_y,
! velocity_mag}_<step:09d>.bin, N_Y rows of N_X values, x fastest).
!
! Coarray images do not exist here: I_X, I_Y and FOR_COARRAY_NUM_IMAGES are
! accepted (domain.f90:73-101) but must describe one image, and the halo
! exchange is part of the device kernel, so the halo timing rows read 0.
program ffb_host
    use iso_fortran_env, only: int32, int64, real32, real64, output_unit, compiler_version, compiler_options
    use iso_c_binding, only: c_ptr, c_null_ptr, c_loc, c_int32_t, c_int64_t, c_double, c_float, c_int
    use ffb_c
    use ffb_settings
    implicit none

    type(run_settings_t) :: s
    type(ffb_config_t) :: cfg
    type(c_ptr) :: lat
    real(c_double), allocatable, target :: rho64(:,:), ux64(:,:), uy64(:,:)
    real(c_float), allocatable, target :: rho32(:,:), ux32(:,:), uy32(:,:)
    logical :: fp64
    integer(int32) :: step, next_stop, chunk
    integer(c_int) :: rc
    integer(int64) :: clock_start, clock_end, clock_rate
    real(real64) :: kernel_seconds, elapsed_seconds, other_seconds, ms, mlups, mass, mx, my, speed
    character(len=:), allocatable :: out_dir
    ! hardware summary (hardware_info.f90:13-29, 132-145): host CPU, threads,
    ! nodes, compiler, flags, commit, plus the GPU the domain lives on
    character(len=:), allocatable :: hw_cpu, hw_compiler, hw_flags, hw_commit, hw_gpu
    integer(int32) :: hw_threads, hw_nodes, hw_cus
    integer(int64) :: hw_gpu_bytes

    call load_settings(s)
    call check_single_domain()
    fp64 = s%precision == "fp64"

    call ffb_config_default(cfg)
    cfg%n_x = s%n_x
    cfg%n_y = s%n_y
    cfg%sim_mode = s%sim_mode
    cfg%omega = s%omega
    cfg%rho_0 = s%rho_0
    cfg%u_max = s%u_max
    cfg%n_sin = s%n_sin
    cfg%u_wall = s%u_wall
    cfg%u_lid = s%u_lid
    cfg%rho_in = s%rho_in
    cfg%rho_out = s%rho_out
    cfg%storage_scale_log2 = s%storage_scale_log2
    cfg%density_checks = merge(1_c_int32_t, 0_c_int32_t, s%density_checks)
    select case (s%storage_rounding)
    case ("stochastic")
        cfg%storage_rounding = FFB_ROUND_STOCHASTIC
    case ("nearest")
        cfg%storage_rounding = FFB_ROUND_NEAREST
    case default
        error stop "error: FFB_STORAGE_ROUNDING must be stochastic or nearest"
    end select
    select case (s%precision)
    case ("fp64")
        cfg%precision = FFB_PREC_FP64
    case ("fp32")
        cfg%precision = FFB_PREC_FP32
    case ("fp16dev")
        cfg%precision = FFB_PREC_FP16_DEV
    case ("bf16dev")
        cfg%precision = FFB_PREC_BF16_DEV
    case ("fp32dev")
        cfg%precision = FFB_PREC_FP32_DEV
    case ("e3m12dev")
        cfg%precision = FFB_PREC_E3M12_DEV
    case default
        error stop "error: FFB_PRECISION must be fp32, fp64, fp16dev, bf16dev, fp32dev or e3m12dev"
    end select
    if (s%storage_scale_log2 == STORAGE_SCALE_DEFAULT) then
        ! the library's default for the form at the mode's speed
        select case (s%sim_mode)
        case (SIM_SHEAR_WAVE)
            speed = s%u_max
        case (SIM_COUETTE_FLOW)
            speed = s%u_wall
        case (SIM_SLIDING_LID)
            speed = s%u_lid
        case default
            speed = max(s%u_lid, s%u_wall, s%u_max)
        end select
        cfg%storage_scale_log2 = ffb_default_scale_log2(cfg%precision, real(abs(speed), c_double))
        if (cfg%storage_scale_log2 == -huge(1_c_int32_t) - 1) &
            error stop "error: no default storage scale for this precision and speed"
    end if

    if (fp64) then
        allocate(rho64(s%n_x, s%n_y), ux64(s%n_x, s%n_y), uy64(s%n_x, s%n_y))
    else
        allocate(rho32(s%n_x, s%n_y), ux32(s%n_x, s%n_y), uy32(s%n_x, s%n_y))
    end if

    call ffb_check(ffb_create(cfg, lat), "ffb_create")
    call ffb_check(ffb_init(lat), "ffb_init")

    call collect_hardware()
    call print_hardware_summary()
    call print_run_summary()
    out_dir = "output/" // trim(s%export_num)
    if (s%export_macros) call export_metadata()
    if (s%export_macros .and. should_export_step(0)) call export_fields(0)

    call print_timestamp("launched")
    kernel_seconds = 0.0_real64
    call system_clock(clock_start, clock_rate)

    step = 0
    do while (step < s%n_steps)
        ! advance to the next step that exports or reports progress
        next_stop = min(s%n_steps, (step / s%progress_interval + 1) * s%progress_interval)
        if (s%export_macros .and. s%export_interval > 0) &
            next_stop = min(next_stop, (step / s%export_interval + 1) * s%export_interval)
        chunk = next_stop - step
        call ffb_check(ffb_reset_timer(lat), "ffb_reset_timer")
        call ffb_check(ffb_step(lat, int(chunk, c_int32_t), &
            merge(1_c_int32_t, 0_c_int32_t, s%export_macros .and. should_export_step(next_stop))), &
            "ffb_step")
        rc = ffb_sync(lat)
        ! FFB_DENSITY_CHECKS: the reference's own stop (simulation.f90:229-233)
        if (rc == FFB_E_DENSITY) error stop ffb_error_message()
        call ffb_check(rc, "ffb_sync")
        call ffb_check(ffb_elapsed_ms(lat, ms), "ffb_elapsed_ms")
        kernel_seconds = kernel_seconds + ms * 1.0e-3_real64
        step = next_stop
        if (s%export_macros .and. should_export_step(step)) call export_fields(step)
        if (mod(step, s%progress_interval) == 0 .or. step == s%n_steps) call print_progress(step)
    end do

    call system_clock(clock_end)
    call print_timestamp("finished")
    elapsed_seconds = real(clock_end - clock_start, real64) / real(clock_rate, real64)
    other_seconds = max(0.0_real64, elapsed_seconds - kernel_seconds)
    if (elapsed_seconds > 0.0_real64) then
        mlups = real(s%n_x, real64) * real(s%n_y, real64) * real(s%n_steps, real64) / elapsed_seconds / 1.0e6_real64
    else
        mlups = 0.0_real64
    end if
    call print_execution_summary()

    call ffb_check(ffb_moments(lat, mass, mx, my), "ffb_moments")
    print '(A,T24,A,ES23.15)', "total mass (FP64 sum)", "= ", mass
    call ffb_check(ffb_destroy(lat), "ffb_destroy")
    flush(output_unit)

contains

    subroutine check_single_domain()
        integer(int32) :: ix, iy, images
        ix = 1
        iy = 1
        images = 1
        call env_count("I_X", ix)
        call env_count("I_Y", iy)
        call env_count("FOR_COARRAY_NUM_IMAGES", images)
        if (ix /= 1 .or. iy /= 1 .or. images /= 1) then
            error stop "error: the device solver runs one domain per GPU; set I_X=1, I_Y=1 " // &
                "and FOR_COARRAY_NUM_IMAGES=1"
        end if
    end subroutine check_single_domain

    subroutine env_count(name, v)
        character(len=*), intent(in) :: name
        integer(int32), intent(inout) :: v
        character(len=32) :: buf
        integer :: n, st, ios
        call get_environment_variable(name, buf, n, st)
        if (st == 0 .and. n > 0) then
            read(buf(1:n), *, iostat=ios) v
            if (ios /= 0) error stop "error: bad value in environment variable " // name
        end if
    end subroutine env_count

    ! should_export_step, source/export.f90:65-89
    logical function should_export_step(st)
        integer(int32), intent(in) :: st
        if (s%export_interval <= 0) then
            should_export_step = .false.
        else if (st == 0) then
            should_export_step = s%export_endpoint_states
        else
            should_export_step = mod(st, s%export_interval) == 0 .or. &
                (s%export_endpoint_states .and. st == s%n_steps)
        end if
    end function should_export_step

    subroutine print_timestamp(what)
        character(len=*), intent(in) :: what
        character(len=10) :: t
        call date_and_time(time=t)
        print '(A)', "[" // t(1:2) // ":" // t(3:4) // ":" // t(5:6) // "] " // what // &
            " ------------------------------------------------------------"
        flush(output_unit)
    end subroutine print_timestamp

    subroutine print_progress(st)
        integer(int32), intent(in) :: st
        integer(int64) :: now
        real(real64) :: el, pct, avg
        character(len=10) :: t
        call system_clock(now)
        call date_and_time(time=t)
        el = real(now - clock_start, real64) / real(clock_rate, real64)
        pct = 100.0_real64 * real(st, real64) / real(max(1, s%n_steps), real64)
        avg = 1000.0_real64 * el / real(max(1, st), real64)
        print '(3A,F6.2,A,I0,A,I0,A,F0.3,A)', "[", t(1:2) // ":" // t(3:4) // ":" // t(5:6), "] ", &
            pct, " %  ", st, "/", s%n_steps, " steps  |  avg step: ", avg, " ms"
        flush(output_unit)
    end subroutine print_progress

    ! host CPU model and logical threads from /proc/cpuinfo, nodes from
    ! SLURM_NNODES (1 without it), the compiler and its options from the
    ! compiler itself, the commit from FFB_GIT_COMMIT at build time, and the
    ! GPU from the HIP runtime
    subroutine collect_hardware()
        integer :: u, ios, k, n, st
        character(len=512) :: line
        character(len=64) :: buf
        character(len=256) :: gname
        hw_cpu = "unknown"
        hw_threads = 0
        open(newunit=u, file="/proc/cpuinfo", status="old", action="read", iostat=ios)
        if (ios == 0) then
            do
                read(u, '(A)', iostat=ios) line
                if (ios /= 0) exit
                if (index(line, "processor") == 1) hw_threads = hw_threads + 1
                if (index(line, "model name") == 1 .and. hw_cpu == "unknown") then
                    k = index(line, ":")
                    if (k > 0) hw_cpu = trim(adjustl(line(k + 1:)))
                end if
            end do
            close(u)
        end if
        hw_nodes = 1
        call get_environment_variable("SLURM_NNODES", buf, n, st)
        if (st == 0 .and. n > 0) then
            read(buf(1:n), *, iostat=ios) k
            if (ios == 0) hw_nodes = k
        end if
        hw_compiler = trim(compiler_version())
        hw_flags = trim(compiler_options())
        if (len(hw_flags) == 0) hw_flags = "none"
#ifdef FFB_GIT_COMMIT
        hw_commit = FFB_GIT_COMMIT
#else
        hw_commit = "unknown"
#endif
        gname = ""
        hw_cus = 0
        hw_gpu_bytes = 0
        if (ffb_device_properties(0_c_int32_t, gname, int(len(gname), c_int32_t), hw_cus, hw_gpu_bytes) == 0) then
            k = index(gname, char(0))
            if (k > 0) gname(k:) = ""
            hw_gpu = trim(gname)
        else
            hw_gpu = "unknown"
        end if
    end subroutine collect_hardware

    ! the reference's block (hardware_info.f90:138-144) plus the GPU line
    subroutine print_hardware_summary()
        character(len=32) :: t
        print '(A)', ""
        print '(A)', "--- [ hardware ] ---------------------------------------------------------------"
        print '(A,T27,A,A)', "cpu model", "= ", hw_cpu
        write(t, '(I0)') hw_threads
        print '(A,T27,A,A)', "logical threads/node", "= ", trim(t)
        write(t, '(I0)') hw_nodes
        print '(A,T27,A,A)', "compute nodes", "= ", trim(t)
        print '(A,T27,A,A)', "compiler", "= ", hw_compiler
        print '(A,T27,A,A)', "flags", "= ", hw_flags
        print '(A,T27,A,A)', "latest commit", "= ", hw_commit
        write(t, '(I0,A,F0.1,A)') hw_cus, " CUs, ", real(hw_gpu_bytes, real64) * 1.0e-9_real64, " GB"
        print '(A,T27,A,A)', "gpu", "= ", hw_gpu // ", " // trim(t) // ", one device-resident domain"
    end subroutine print_hardware_summary

    subroutine print_run_summary()
        real(real64) :: bytes_pop, gb
        bytes_pop = merge(8.0_real64, merge(4.0_real64, 2.0_real64, &
                          s%precision == "fp32" .or. s%precision == "fp32dev"), fp64)
        gb = 2.0_real64 * 9.0_real64 * bytes_pop * real(s%n_x, real64) * real(s%n_y, real64) * 1.0e-9_real64
        print '(A)', ""
        print '(A)', "--- [ simulation parameters ] --------------------------------------------------"
        print '(A,T27,A,A)', "SIM_MODE", "= ", mode_name(s%sim_mode)
        print '(A,T27,A,F8.6)', "rho_0", "= ", s%rho_0
        print '(A,T27,A,F8.6)', "omega", "= ", s%omega
        select case (s%sim_mode)
        case (SIM_SHEAR_WAVE)
            print '(A,T27,A,F8.6)', "u_max", "= ", s%u_max
            print '(A,T27,A,F8.6)', "n_sin", "= ", s%n_sin
        case (SIM_COUETTE_FLOW)
            print '(A,T27,A,F8.6)', "u_wall", "= ", s%u_wall
            print '(A,T27,A,F0.3)', "reynolds", "= ", reynolds_number(s)
        case (SIM_POISEUILLE_FLOW)
            print '(A,T27,A,F8.6)', "rho_in", "= ", s%rho_in
            print '(A,T27,A,F8.6)', "rho_out", "= ", s%rho_out
        case (SIM_SLIDING_LID)
            print '(A,T27,A,F8.6)', "u_lid", "= ", s%u_lid
            print '(A,T27,A,F0.3)', "reynolds", "= ", reynolds_number(s)
        end select
        print '(A,T27,A,I0,A,I0,A,I0)', "grid [X/Y/N]", "= ", s%n_x, " x ", s%n_y, " x ", s%n_steps
        print '(A,T27,A,A)', "precision", "= ", s%precision
        print '(A,T27,A,F0.3,A)', "population buffers", "= ", gb, " GB"
        print '(A)', ""
        flush(output_unit)
    end subroutine print_run_summary

    subroutine print_execution_summary()
        print '(A)', ""
        print '(A)', "image execution time spread  |      best [sec] (image) |     worst [sec] (image)"
        print '(A)', "--------------------------------------------------------------------------------"
        call spread_row("kernel compute", kernel_seconds)
        call spread_row("halo sync", 0.0_real64)
        call spread_row("halo transfer", 0.0_real64)
        call spread_row("other", other_seconds)
        call spread_row("total", elapsed_seconds)
        print '(A)', ""
        print '(A)', "--- [ perf metrics ] -----------------------------------------------------------"
        print '(A,T24,A,I0,A,I0,A,I0,A)', "sim size [X/Y/N]", "= [ ", s%n_x, " / ", s%n_y, " / ", s%n_steps, " ]"
        print '(A,T24,A,A,A)', "total time", "= ", compact(elapsed_seconds), " sec"
        print '(A,T24,A,A,A)', "step time", "= ", &
            compact(1000.0_real64 * elapsed_seconds / real(max(1, s%n_steps), real64)), " ms"
        print '(A,T24,A,I0)', "MLUPS", "= ", int(mlups, int64)
    end subroutine print_execution_summary

    subroutine spread_row(name, seconds)
        character(len=*), intent(in) :: name
        real(real64), intent(in) :: seconds
        character(len=28) :: row
        character(len=23) :: cell
        character(len=64) :: raw
        write(raw, '(F0.3," (",I0,")")') seconds, 1
        raw = adjustl(raw)
        if (raw(1:1) == ".") raw = "0" // raw
        cell = repeat(" ", 23)
        cell(24 - len_trim(raw):) = trim(raw)
        row = name
        print '(A)', row // " | " // cell // " | " // cell
    end subroutine spread_row

    function compact(v) result(txt)
        real(real64), intent(in) :: v
        character(len=:), allocatable :: txt
        character(len=32) :: raw
        write(raw, '(F0.3)') v
        raw = adjustl(raw)
        if (raw(1:1) == ".") then
            txt = "0" // trim(raw)
        else
            txt = trim(raw)
        end if
    end function compact

    subroutine export_metadata()
        integer :: u
        call execute_command_line("mkdir -p " // out_dir)
        open(newunit=u, file=out_dir // "/config.json", form="formatted", status="replace", action="write")
        write(u, '(A)') "{"
        ! hardware block (hardware_info.f90:257-272), plus the GPU
        write(u, '(A)') '  "hardware": {'
        write(u, '(A,A,A)') '    "cpu_model": "', json_escape(hw_cpu), '",'
        write(u, '(A,I0,A)') '    "logical_threads_per_node": ', hw_threads, ','
        write(u, '(A,I0,A)') '    "compute_nodes": ', hw_nodes, ','
        write(u, '(A,A,A)') '    "compiler": "', json_escape(hw_compiler), '",'
        write(u, '(A,A,A)') '    "compiler_flags": "', json_escape(hw_flags), '",'
        write(u, '(A,A,A)') '    "gpu": "', json_escape(hw_gpu), '",'
        write(u, '(A,I0,A)') '    "gpu_compute_units": ', hw_cus, ','
        write(u, '(A,ES24.16)') '    "gpu_memory_GB": ', real(hw_gpu_bytes, real64) * 1.0e-9_real64
        write(u, '(A)') '  },'
        write(u, '(A)') '  "device": "' // json_escape(hw_gpu) // '",'
        write(u, '(A,A,A)') '  "SIM_MODE": "', mode_name(s%sim_mode), '",'
        write(u, '(A,ES24.16,A)') '  "rho_0": ', s%rho_0, ','
        write(u, '(A,ES24.16,A)') '  "omega": ', s%omega, ','
        select case (s%sim_mode)
        case (SIM_SHEAR_WAVE)
            write(u, '(A,ES24.16,A)') '  "u_max": ', s%u_max, ','
            write(u, '(A,ES24.16,A)') '  "n_sin": ', s%n_sin, ','
        case (SIM_COUETTE_FLOW)
            write(u, '(A,ES24.16,A)') '  "u_wall": ', s%u_wall, ','
            write(u, '(A,ES24.16,A)') '  "reynolds": ', reynolds_number(s), ','
        case (SIM_POISEUILLE_FLOW)
            write(u, '(A,ES24.16,A)') '  "rho_in": ', s%rho_in, ','
            write(u, '(A,ES24.16,A)') '  "rho_out": ', s%rho_out, ','
        case (SIM_SLIDING_LID)
            write(u, '(A,ES24.16,A)') '  "u_lid": ', s%u_lid, ','
            write(u, '(A,ES24.16,A)') '  "reynolds": ', reynolds_number(s), ','
        end select
        write(u, '(A,I0,A)') '  "N_X": ', s%n_x, ','
        write(u, '(A,I0,A)') '  "N_Y": ', s%n_y, ','
        write(u, '(A,I0,A)') '  "N_STEPS": ', s%n_steps, ','
        write(u, '(A,I0,A)') '  "N_CELLS": ', int(s%n_x, int64) * int(s%n_y, int64), ','
        write(u, '(A)') '  "N_DIRS": 9,'
        write(u, '(A)') '  "dist_func_layout": "xyd",'
        write(u, '(A)') '  "use_unrolled_kernels": true,'
        write(u, '(A,A,A)') '  "FFB_FP64": ', trim(merge("true ", "false", fp64)), ','
        write(u, '(A,A,A)') '  "FFB_DENSITY_CHECKS": ', trim(merge("true ", "false", s%density_checks)), ','
        write(u, '(A,A,A)') '  "precision": "', s%precision, '",'
        write(u, '(A,A,A)') '  "export_macros": ', merge("true ", "false", s%export_macros), ','
        write(u, '(A,A,A)') '  "export_endpoint_states": ', merge("true ", "false", s%export_endpoint_states), ','
        write(u, '(A,I0,A)') '  "export_interval": ', s%export_interval, ','
        write(u, '(A)') '  "domain_decomposition": {'
        write(u, '(A)') '    "coarray_images": 1,'
        write(u, '(A)') '    "image_grid_x": 1,'
        write(u, '(A)') '    "image_grid_y": 1,'
        write(u, '(A,I0,A)') '    "local_n_x": ', s%n_x, ','
        write(u, '(A,I0,A)') '    "local_n_y": ', s%n_y, ','
        ! the device domain holds no halo cells (boundary values are computed
        ! where they are read); the reference's layout holds
        ! (n_x + 2)(n_y + 2) - n_x n_y of them (export.f90:154-157)
        write(u, '(A)') '    "halo_cells_percent": 0.0'
        write(u, '(A)') '  },'
        call write_memory_usage(u)
        write(u, '(A,A,A)') '  "output_dir": "', out_dir, '",'
        write(u, '(A,A,A)') '  "file_dtype": "', merge("real64", "real32", fp64), '"'
        write(u, '(A)') "}"
        close(u)
    end subroutine export_metadata

    ! memory_usage block (export.f90:237-247): device population sets and macro
    ! fields, plus the host arrays the export reads the fields into
    subroutine write_memory_usage(u)
        integer, intent(in) :: u
        integer(c_int64_t) :: pop_b, mac_b, oth_b
        real(real64) :: cells, host_b, total
        call ffb_check(ffb_memory_usage(lat, pop_b, mac_b, oth_b), "ffb_memory_usage")
        cells = real(s%n_x, real64) * real(s%n_y, real64)
        host_b = 3.0_real64 * cells * merge(8.0_real64, 4.0_real64, fp64)
        total = real(pop_b, real64) + real(mac_b, real64) + real(oth_b, real64) + host_b
        write(u, '(A)') '  "memory_usage": {'
        write(u, '(A,I0,A)') '    "dist_function_buffers_per_cell_B": ', nint(real(pop_b, real64) / cells, int64), ','
        write(u, '(A,ES24.16,A)') '    "dist_function_buffers_total_GB": ', real(pop_b, real64) * 1.0e-9_real64, ','
        write(u, '(A,I0,A)') '    "macro_field_buffers_per_cell_B": ', &
            nint((real(mac_b, real64) + host_b) / cells, int64), ','
        write(u, '(A,ES24.16,A)') '    "macro_field_buffers_total_GB": ', &
            (real(mac_b, real64) + host_b) * 1.0e-9_real64, ','
        write(u, '(A,I0,A)') '    "total_per_cell_B": ', nint(total / cells, int64), ','
        write(u, '(A,ES24.16)') '    "total_GB": ', total * 1.0e-9_real64
        write(u, '(A)') '  },'
    end subroutine write_memory_usage

    ! JSON string escape of backslash and double quote
    function json_escape(v) result(o)
        character(len=*), intent(in) :: v
        character(len=:), allocatable :: o
        integer :: k
        o = ""
        do k = 1, len_trim(v)
            select case (v(k:k))
            case ('"', '\')
                o = o // '\' // v(k:k)
            case default
                o = o // v(k:k)
            end select
        end do
    end function json_escape

    subroutine export_fields(st)
        integer(int32), intent(in) :: st
        character(len=10) :: suffix
        write(suffix, '("_",I9.9)') st
        if (fp64) then
            call ffb_check(ffb_read_macros(lat, c_loc(rho64), c_loc(ux64), c_loc(uy64)), "ffb_read_macros")
            call write_field64(out_dir // "/density" // suffix // ".bin", rho64)
            call write_field64(out_dir // "/velocity_x" // suffix // ".bin", ux64)
            call write_field64(out_dir // "/velocity_y" // suffix // ".bin", uy64)
            call write_field64(out_dir // "/velocity_mag" // suffix // ".bin", sqrt(ux64 * ux64 + uy64 * uy64))
        else
            call ffb_check(ffb_read_macros(lat, c_loc(rho32), c_loc(ux32), c_loc(uy32)), "ffb_read_macros")
            call write_field32(out_dir // "/density" // suffix // ".bin", rho32)
            call write_field32(out_dir // "/velocity_x" // suffix // ".bin", ux32)
            call write_field32(out_dir // "/velocity_y" // suffix // ".bin", uy32)
            call write_field32(out_dir // "/velocity_mag" // suffix // ".bin", sqrt(ux32 * ux32 + uy32 * uy32))
        end if
    end subroutine export_fields

    ! one stream-unformatted file, N_Y rows of N_X values (export.f90:496-521)
    subroutine write_field64(path, field)
        character(len=*), intent(in) :: path
        real(real64), intent(in) :: field(:,:)
        integer :: u
        open(newunit=u, file=path, access="stream", form="unformatted", status="replace", action="write")
        write(u) field
        close(u)
    end subroutine write_field64

    subroutine write_field32(path, field)
        character(len=*), intent(in) :: path
        real(real32), intent(in) :: field(:,:)
        integer :: u
        open(newunit=u, file=path, access="stream", form="unformatted", status="replace", action="write")
        write(u) field
        close(u)
    end subroutine write_field32

end program ffb_host
