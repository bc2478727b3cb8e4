! mfcfd_dropin.f90 -- the reference's hot-path modules, computed on the GPU.
!
! Same module and procedure names, same argument lists and the same
! data_structure_mod globals as the reference's files, so the reference's
! own q_lskum (q_lskum.f90:15-36), meshfree_solver and post_processor link
! against these unchanged:
!   compute_conserved_vector_mod      compute_conserved_vector.f90:13
!   q_variables_mod                   q_variables.f90:10
!   q_derivatives_mod                 q_derivatives.f90:10
!   timestep_delt_mod                 timestep_delt.f90:8
!   flux_residual_mod                 flux_residual.f90:30
!   implicit_forward_and_reverse_sweeps_mod
!                                     implicit_forward_and_reverse_sweeps.f90:16
!   state_update_mod                  state_update.f90:15
!   fpi_solver_mod                    fpi_solver.f90:20
!
! fpi_solver keeps the flow state in device memory between calls and updates
! res_new, residue, res_old, max_res and max_res_point after every call; the
! host arrays point%prim, U, q, qm, dq, delt, flux_res, delUp and delUn are
! refreshed after the call with t = max_iters (so post-processing reads the
! final state), whenever a standalone entry point is called, or on
! mfcfd_sync_host().  The host must not change point%... between fpi_solver
! calls (q_lskum does not).  A standalone entry point reads its inputs from
! point%... and writes its outputs back, like the reference routine.

module compute_conserved_vector_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine compute_conserved_vector()
        call mfcfd_entry_inputs([ARR_PRIM])
        call mfcfd_check(mfcfd_compute_conserved_vector(mfcfd_ctx), 'compute_conserved_vector')
        call mfcfd_entry_outputs([ARR_U])
    end subroutine
end module compute_conserved_vector_mod

module q_variables_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine eval_q_variables()
        call mfcfd_entry_inputs([ARR_PRIM])
        call mfcfd_check(mfcfd_eval_q_variables(mfcfd_ctx), 'eval_q_variables')
        call mfcfd_entry_outputs([ARR_Q])
    end subroutine
end module q_variables_mod

module q_derivatives_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine eval_q_derivatives()
        call mfcfd_entry_inputs([ARR_Q])
        call mfcfd_check(mfcfd_eval_q_derivatives(mfcfd_ctx), 'eval_q_derivatives')
        call mfcfd_entry_outputs([ARR_QM, ARR_DQ])
    end subroutine
end module q_derivatives_mod

module timestep_delt_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine timestep_delt()
        call mfcfd_entry_inputs([ARR_PRIM])
        call mfcfd_check(mfcfd_timestep_delt(mfcfd_ctx), 'timestep_delt')
        call mfcfd_entry_outputs([ARR_DELT])
    end subroutine
end module timestep_delt_mod

module flux_residual_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine eval_flux_residual()
        call mfcfd_entry_inputs([ARR_Q, ARR_QM, ARR_DQ])
        call mfcfd_check(mfcfd_eval_flux_residual(mfcfd_ctx), 'eval_flux_residual')
        call mfcfd_entry_outputs([ARR_FLUX_RES])
    end subroutine
end module flux_residual_mod

module implicit_forward_and_reverse_sweeps_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine implicit_forward_backward_sweeps()
        call mfcfd_entry_inputs([ARR_PRIM, ARR_U, ARR_DELT, ARR_FLUX_RES, ARR_DELUP, ARR_DELUN])
        call mfcfd_check(mfcfd_implicit_forward_backward_sweeps(mfcfd_ctx), &
                         'implicit_forward_backward_sweeps')
        call mfcfd_entry_outputs([ARR_DELUP, ARR_DELUN])
    end subroutine
end module implicit_forward_and_reverse_sweeps_mod

module state_update_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine state_update()
        type(mfcfd_residual_t) :: r
        call mfcfd_entry_inputs([ARR_PRIM, ARR_U, ARR_DELT, ARR_FLUX_RES, ARR_DELUN])
        call mfcfd_check(mfcfd_state_update(mfcfd_ctx), 'state_update')
        call mfcfd_entry_outputs([ARR_PRIM, ARR_DELUP])
        call mfcfd_check(mfcfd_get_residual(mfcfd_ctx, r), 'get_residual')
        res_new = r%res_new
        max_res = r%max_res
        if (r%max_res_point > 0) max_res_point = r%max_res_point
    end subroutine
end module state_update_mod

module fpi_solver_mod
    use data_structure_mod
    use mfcfd_device_mod
    implicit none
contains
    subroutine fpi_solver(t)
        integer :: t
        type(mfcfd_residual_t) :: r
        real(c_double) :: rn, rs
        call mfcfd_ensure_bound()
        if (.not. mfcfd_device_ahead) then
            ! the host holds the current state (first call, or after a
            ! standalone entry point or a sync)
            call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_PRIM, point%prim), 'upload prim')
            call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELUP, point%delUp), 'upload delUp')
            call mfcfd_check(mfcfd_upload(mfcfd_ctx, ARR_DELUN, point%delUn), 'upload delUn')
        endif
        call mfcfd_check(mfcfd_fpi_solver(mfcfd_ctx, int(t, c_int), rn, rs), 'fpi_solver')
        mfcfd_device_ahead = .true.
        call mfcfd_check(mfcfd_get_residual(mfcfd_ctx, r), 'get_residual')
        res_new = r%res_new
        residue = r%residue
        res_old = r%res_old
        max_res = r%max_res
        if (r%max_res_point > 0) max_res_point = r%max_res_point
        if (t >= max_iters) call mfcfd_sync_host()
    end subroutine
end module fpi_solver_mod
