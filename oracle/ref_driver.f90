! ref_driver.f90 -- dump driver for the reference Fortran build (oracle/_ref).
!
! TEST INFRASTRUCTURE ONLY.  Linked against the reference's own module files
! (compiled from their original location by oracle/build_ref.sh), it calls
! the public module procedures in the order q_lskum.f90:24-32 and
! fpi_solver.f90:27-43 use and writes raw little-endian binary dumps:
!
!   pre.bin    split-stencil counts and lists, alias, point_with_alias
!   hist.bin   per iteration: res_new, residue, max_res (float64)
!   start.bin  state before the last iteration (prim, delUp, delUn, res_old)
!   step.bin   the last iteration, entry point by entry point:
!              U, q, qm, dq, delt, flux_res, delUp, delUn, prim, delUp, res_new
!   final.bin  prim, U, q, qm, dq, delt, flux_res, delUp, delUn
!   time.txt   seconds spent in the timed fpi_solver loop, iterations timed
!   trace.bin  (trace = 1 only) per iteration: the largest |p - pr_inf| over
!              the outer points, the smallest pressure over all points and
!              the number of points with a non-finite pressure (float64)
!
! Usage (in a directory holding 3d_input_data and stored-solution.dat):
!   ref_driver <iterations> <outdir> [stepwise(0|1) [trace(0|1)]]
program ref_driver
    use data_structure_mod
    use point_preprocessor_mod
    use initial_conditions_mod
    use generate_connectivity_mod
    use implicit_aliasing_mod
    use fpi_solver_mod
    use compute_conserved_vector_mod
    use q_variables_mod
    use q_derivatives_mod
    use timestep_delt_mod
    use flux_residual_mod
    use implicit_forward_and_reverse_sweeps_mod
    use state_update_mod
    implicit none
    integer :: t, niter, stepwise, ntimed, u, trace, m, k
    integer(kind=8) :: c0, c1, crate
    character(len=512) :: arg, outdir
    real*8, allocatable :: hist(:,:), tr(:,:)
    real*8 :: pk

    call get_command_argument(1, arg)
    read(arg, *) niter
    call get_command_argument(2, outdir)
    stepwise = 0
    if (command_argument_count() >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) stepwise
    endif
    trace = 0
    if (command_argument_count() >= 4) then
        call get_command_argument(4, arg)
        read(arg, *) trace
    endif
    allocate(hist(3, max(niter, 1)), tr(3, max(niter, 1)))
    hist = 0.0d0
    tr = 0.0d0

    call allocate_size()
    call read_input_point_data()
    call initial_conditions()
    call generate_split_stencils()
    call aliasing()

    open(newunit=u, file=trim(outdir)//'/pre.bin', form='unformatted', access='stream', status='replace')
    write(u) point%xpos_nbhs, point%xneg_nbhs, point%ypos_nbhs, point%yneg_nbhs, &
             point%zpos_nbhs, point%zneg_nbhs
    write(u) point%xpos_conn, point%xneg_conn, point%ypos_conn, point%yneg_conn, &
             point%zpos_conn, point%zneg_conn
    write(u) point%alias, point%point_with_alias
    close(u)

    ntimed = niter
    if (stepwise /= 0) ntimed = niter - 1
    call system_clock(c0, crate)
    do t = 1, ntimed
        call fpi_solver(t)
        hist(1, t) = res_new
        hist(2, t) = residue
        hist(3, t) = max_res
        if (trace /= 0) then
            do m = 1, outer_points
                k = outer_points_index(m)
                tr(1, t) = max(tr(1, t), abs(point%prim(5, k) - pr_inf))
            enddo
            tr(2, t) = minval(point%prim(5, :))
            do k = 1, max_points
                pk = point%prim(5, k)
                if (.not. (pk == pk) .or. abs(pk) > huge(pk)) tr(3, t) = tr(3, t) + 1.0d0
            enddo
        endif
    enddo
    call system_clock(c1)
    open(newunit=u, file=trim(outdir)//'/time.txt', status='replace')
    write(u, *) dble(c1 - c0) / dble(crate), ntimed
    close(u)

    if (stepwise /= 0 .and. niter >= 1) then
        t = niter
        open(newunit=u, file=trim(outdir)//'/start.bin', form='unformatted', access='stream', status='replace')
        write(u) point%prim, point%delUp, point%delUn, res_old
        close(u)
        open(newunit=u, file=trim(outdir)//'/step.bin', form='unformatted', access='stream', status='replace')
        call compute_conserved_vector()
        write(u) point%U
        call eval_q_variables()
        write(u) point%q
        call eval_q_derivatives()
        write(u) point%qm, point%dq
        call timestep_delt()
        write(u) point%delt
        call eval_flux_residual()
        write(u) point%flux_res
        call implicit_forward_backward_sweeps()
        write(u) point%delUp, point%delUn
        call state_update()
        write(u) point%prim, point%delUp, res_new
        close(u)
        if (t .le. 2) then
            res_old = res_new
            residue = 0.0d0
        else
            residue = dlog10(res_new/res_old)
        endif
        hist(1, t) = res_new
        hist(2, t) = residue
        hist(3, t) = max_res
    endif

    if (trace /= 0) then
        open(newunit=u, file=trim(outdir)//'/trace.bin', form='unformatted', access='stream', status='replace')
        write(u) tr(:, 1:niter)
        close(u)
    endif
    open(newunit=u, file=trim(outdir)//'/hist.bin', form='unformatted', access='stream', status='replace')
    write(u) hist(:, 1:niter)
    close(u)
    open(newunit=u, file=trim(outdir)//'/final.bin', form='unformatted', access='stream', status='replace')
    write(u) point%prim, point%U, point%q, point%qm, point%dq, point%delt, point%flux_res, &
             point%delUp, point%delUn
    close(u)
end program ref_driver
