! ref_explicit_sweep.f90 -- explicit-mode stand-in for the sweep module.
!
! TEST INFRASTRUCTURE ONLY.  The reference keeps its explicit update only as
! commented lines (state_update.f90:35,67: U = U - delt*flux_res).  Linking
! this module in place of implicit_forward_and_reverse_sweeps.f90 gives that
! update through the live state_update.f90: delUn = -delt*flux_res, then
! delUp = delUn as the implicit sweep ends (implicit_forward_and_reverse_
! sweeps.f90:86-88).  U + (-(delt*R)) rounds exactly like U - delt*R.
module implicit_forward_and_reverse_sweeps_mod
    use data_structure_mod
    implicit none
contains
    subroutine implicit_forward_backward_sweeps()
        integer :: k
        do k = 1, max_points
            point%delUn(:,k) = -point%delt(k)*point%flux_res(:,k)
            point%delUp(:,k) = point%delUn(:,k)
        enddo
    end subroutine
end module implicit_forward_and_reverse_sweeps_mod
