! mfcfd_postprocess.f90 -- post_processor_mod for the drop-in driver.
!
! print_primal_output writes the files of post_processor.f90:13-45 with the
! same list-directed records (x, y, z, prim(1:5) per point in
! primal-solution.dat; k, status, flux_res(1:5) in flux-residual.dat;
! k, status, dq(3,1:5) in qderivatives.dat; normals.dat empty), from the
! host arrays after the device state has been copied back.
!
! The reference's routine opens unit 101 with ACTION="WRITE" while
! point_preprocessor.f90:25 and initial_conditions.f90:15 left it connected
! with ACTION="READ"; the flang runtime rejects that re-open
! (post_processor.f90:20), so the reference program ends in a runtime error
! before writing its solution.  This version closes units 101-104 first.
module post_processor_mod
    use data_structure_mod
    use mfcfd_device_mod, only: mfcfd_device_ahead, mfcfd_sync_host
    implicit none
contains

    subroutine print_primal_output()
        integer :: k, r, u
        logical :: op
        if (mfcfd_device_ahead) call mfcfd_sync_host()
        do u = 101, 104
            inquire(unit=u, opened=op)
            if (op) close(u)
        end do
        open(unit=101, file="primal-solution.dat", form="FORMATTED", status="REPLACE", action="WRITE")
        open(unit=102, file="normals.dat", form="FORMATTED", status="REPLACE", action="WRITE")
        open(unit=103, file="flux-residual.dat", form="FORMATTED", status="REPLACE", action="WRITE")
        open(unit=104, file="qderivatives.dat", form="FORMATTED", status="REPLACE", action="WRITE")
        do k = 1, max_points
            write(101, *) point%x(k), point%y(k), point%z(k), (point%prim(r, k), r = 1, 5)
            write(104, *) k, point%status(k), (point%dq(3, r, k), r = 1, 5)
            write(103, *) k, point%status(k), (point%flux_res(r, k), r = 1, 5)
        end do
        do u = 101, 104
            close(u)
        end do
    end subroutine

end module post_processor_mod
