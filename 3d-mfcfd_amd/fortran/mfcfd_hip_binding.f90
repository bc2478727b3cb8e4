! mfcfd_hip_binding.f90 -- iso_c_binding interfaces to libmfcfd_hip.so
! (include/mfcfd_hip.h) and the device context shared by the drop-in
! modules in mfcfd_dropin.f90.
!
! The device context is created and bound on first use from the module
! globals of data_structure_mod and the PARAMETERs of parameter_mod
! (data_structure.f90:46-59, parameter.f90:8-33).  The iteration form is the
! reference's canonical implicit one unless the environment variable
! MFCFD_ITERATION_FORM is "explicit" (U <- U - delt*flux_res,
! state_update.f90:35,67).
module mfcfd_hip_binding
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: MFCFD_EXPLICIT = 0, MFCFD_IMPLICIT = 1, MFCFD_IMPLICIT_BACKUP = 2
    integer(c_int), parameter :: MFCFD_FP64 = 0, MFCFD_FLUX_FUSED = 0, MFCFD_SWEEP_ASYNC = 0
    integer(c_int), parameter :: ARR_PRIM = 0, ARR_U = 1, ARR_Q = 2, ARR_QM = 3, ARR_DQ = 4, &
                                 ARR_DELT = 5, ARR_FLUX_RES = 6, ARR_DELUP = 7, ARR_DELUN = 8

    type, bind(C) :: mfcfd_params_t
        real(c_double) :: mach, aoa, cfl, vl_const, power
        integer(c_int) :: inner_iterations, mode, precision, max_iters, flux_order, sweep_schedule
    end type

    type, bind(C) :: mfcfd_residual_t
        real(c_double) :: res_old, res_new, residue, max_res, sum_res_sqr
        integer(c_int) :: max_res_point, iteration
    end type

    interface
        integer(c_int) function mfcfd_create(device, ctx) bind(C, name="mfcfd_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr) :: ctx
        end function
        integer(c_int) function mfcfd_destroy(ctx) bind(C, name="mfcfd_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function mfcfd_last_error() bind(C, name="mfcfd_last_error")
            import :: c_ptr
        end function
        integer(c_int) function mfcfd_set_params(ctx, p) bind(C, name="mfcfd_set_params")
            import :: c_int, c_ptr, mfcfd_params_t
            type(c_ptr), value :: ctx
            type(mfcfd_params_t), intent(in) :: p
        end function
        integer(c_int) function mfcfd_bind_geometry_arrays(ctx, n, x, y, z, tan1, tan2, nor, &
                min_dist, status, nbhs, conn, xpos_nbhs, xneg_nbhs, ypos_nbhs, yneg_nbhs, &
                zpos_nbhs, zneg_nbhs, xpos_conn, xneg_conn, ypos_conn, yneg_conn, zpos_conn, &
                zneg_conn, alias, point_with_alias) bind(C, name="mfcfd_bind_geometry_arrays")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
            real(c_double), intent(in) :: x(*), y(*), z(*), tan1(*), tan2(*), nor(*), min_dist(*)
            integer(c_int), intent(in) :: status(*), nbhs(*), conn(*)
            integer(c_int), intent(in) :: xpos_nbhs(*), xneg_nbhs(*), ypos_nbhs(*), yneg_nbhs(*)
            integer(c_int), intent(in) :: zpos_nbhs(*), zneg_nbhs(*)
            integer(c_int), intent(in) :: xpos_conn(*), xneg_conn(*), ypos_conn(*), yneg_conn(*)
            integer(c_int), intent(in) :: zpos_conn(*), zneg_conn(*)
            integer(c_int), intent(in) :: alias(*), point_with_alias(*)
        end function
        integer(c_int) function mfcfd_upload(ctx, id, host) bind(C, name="mfcfd_upload")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: id
            real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function mfcfd_download(ctx, id, host) bind(C, name="mfcfd_download")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: id
            real(c_double), intent(inout) :: host(*)
        end function
        integer(c_int) function mfcfd_compute_conserved_vector(ctx) &
                bind(C, name="mfcfd_compute_conserved_vector")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_eval_q_variables(ctx) bind(C, name="mfcfd_eval_q_variables")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_eval_q_derivatives(ctx) bind(C, name="mfcfd_eval_q_derivatives")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_timestep_delt(ctx) bind(C, name="mfcfd_timestep_delt")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_eval_flux_residual(ctx) bind(C, name="mfcfd_eval_flux_residual")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_implicit_forward_backward_sweeps(ctx) &
                bind(C, name="mfcfd_implicit_forward_backward_sweeps")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_state_update(ctx) bind(C, name="mfcfd_state_update")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfcfd_fpi_solver(ctx, t, res_new, residue) bind(C, name="mfcfd_fpi_solver")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: t
            real(c_double) :: res_new, residue
        end function
        integer(c_int) function mfcfd_get_residual(ctx, r) bind(C, name="mfcfd_get_residual")
            import :: c_int, c_ptr, mfcfd_residual_t
            type(c_ptr), value :: ctx
            type(mfcfd_residual_t) :: r
        end function
    end interface

contains

    subroutine mfcfd_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (rc == 0) return
        call c_f_pointer(mfcfd_last_error(), msg, [512])
        text = ' '
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        enddo
        write(0, '(a, a, a, i0, a, a)') 'mfcfd: ', what, ' failed (', rc, '): ', trim(text)
        error stop 'mfcfd device call failed'
    end subroutine

end module mfcfd_hip_binding


! The device context of the drop-in modules and its coherence with the host
! arrays of data_structure_mod.
module mfcfd_device_mod
    use iso_c_binding
    use data_structure_mod
    use mfcfd_hip_binding
    implicit none

    type(c_ptr), save :: mfcfd_ctx = c_null_ptr
    ! .true. while the device state is newer than point%(prim, U, ...)
    logical, save :: mfcfd_device_ahead = .false.

contains

    ! create the context, set parameter_mod's values and bind the point arrays
    subroutine mfcfd_ensure_bound()
        type(mfcfd_params_t) :: p
        character(len=32) :: form
        integer :: stat
        if (c_associated(mfcfd_ctx)) return
        call mfcfd_check(mfcfd_create(0_c_int, mfcfd_ctx), 'mfcfd_create')
        p%mach = Mach
        p%aoa = aoa
        p%cfl = CFL
        p%vl_const = VL_CONST
        p%power = power
        p%inner_iterations = inner_iterations
        p%mode = MFCFD_IMPLICIT
        call get_environment_variable('MFCFD_ITERATION_FORM', form, status=stat)
        if (stat == 0 .and. trim(form) == 'explicit') p%mode = MFCFD_EXPLICIT
        if (stat == 0 .and. trim(form) == 'implicit_backup') p%mode = MFCFD_IMPLICIT_BACKUP
        p%precision = MFCFD_FP64
        p%max_iters = max_iters
        p%flux_order = MFCFD_FLUX_FUSED
        p%sweep_schedule = MFCFD_SWEEP_ASYNC
        call mfcfd_check(mfcfd_set_params(mfcfd_ctx, p), 'mfcfd_set_params')
        call mfcfd_check(mfcfd_bind_geometry_arrays(mfcfd_ctx, int(max_points, c_int), point%x, &
            point%y, point%z, point%tan1, point%tan2, point%nor, point%min_dist, point%status, &
            point%nbhs, point%conn, point%xpos_nbhs, point%xneg_nbhs, point%ypos_nbhs, &
            point%yneg_nbhs, point%zpos_nbhs, point%zneg_nbhs, point%xpos_conn, point%xneg_conn, &
            point%ypos_conn, point%yneg_conn, point%zpos_conn, point%zneg_conn, point%alias, &
            point%point_with_alias), 'mfcfd_bind_geometry')
        call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_PRIM, point%prim), 'upload prim')
        call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELUP, point%delUp), 'upload delUp')
        call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELUN, point%delUn), 'upload delUn')
        mfcfd_device_ahead = .false.
    end subroutine

    ! copy every state array the hot path writes back into point%...
    subroutine mfcfd_sync_host()
        if (.not. c_associated(mfcfd_ctx)) return
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_PRIM, point%prim), 'download prim')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_U, point%U), 'download U')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_Q, point%q), 'download q')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_QM, point%qm), 'download qm')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DQ, point%dq), 'download dq')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DELT, point%delt), 'download delt')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_FLUX_RES, point%flux_res), 'download flux_res')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DELUP, point%delUp), 'download delUp')
        call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DELUN, point%delUn), 'download delUn')
        mfcfd_device_ahead = .false.
    end subroutine

    ! Standalone entry point: make the host current, then send the arrays
    ! the routine reads.
    subroutine mfcfd_entry_inputs(ids)
        integer(c_int), intent(in) :: ids(:)
        integer :: m
        call mfcfd_ensure_bound()
        if (mfcfd_device_ahead) call mfcfd_sync_host()
        do m = 1, size(ids)
            select case (ids(m))
            case (ARR_PRIM); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_PRIM, point%prim), 'upload prim')
            case (ARR_U); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_U, point%U), 'upload U')
            case (ARR_Q); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_Q, point%q), 'upload q')
            case (ARR_QM); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_QM, point%qm), 'upload qm')
            case (ARR_DQ); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DQ, point%dq), 'upload dq')
            case (ARR_DELT); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELT, point%delt), 'upload delt')
            case (ARR_FLUX_RES)
                call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_FLUX_RES, point%flux_res), 'upload flux_res')
            case (ARR_DELUP); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELUP, point%delUp), 'upload delUp')
            case (ARR_DELUN); call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELUN, point%delUn), 'upload delUn')
            end select
        enddo
    end subroutine

    ! Standalone entry point: bring back what the routine wrote.
    subroutine mfcfd_entry_outputs(ids)
        integer(c_int), intent(in) :: ids(:)
        integer :: m
        do m = 1, size(ids)
            select case (ids(m))
            case (ARR_PRIM); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_PRIM, point%prim), 'download prim')
            case (ARR_U); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_U, point%U), 'download U')
            case (ARR_Q); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_Q, point%q), 'download q')
            case (ARR_QM); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_QM, point%qm), 'download qm')
            case (ARR_DQ); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DQ, point%dq), 'download dq')
            case (ARR_DELT); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DELT, point%delt), 'download delt')
            case (ARR_FLUX_RES)
                call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_FLUX_RES, point%flux_res), 'download flux_res')
            case (ARR_DELUP); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DELUP, point%delUp), 'download delUp')
            case (ARR_DELUN); call mfcfd_check(mfcfd_download(mfcfd_ctx, ARR_DELUN, point%delUn), 'download delUn')
            end select
        enddo
    end subroutine

end module mfcfd_device_mod
