!-----------------------------------------------------------------------------------------------
! Drop-in bodies for the reference's force / neighbour / integration routines.
!
! Link these in place of the reference's own definitions (src/linklist.f90 assign_to_cell,
! assign_to_neighbourlist, linklist_deallocateall, sort_mols; src/simulation_compute_forces.f90
! simulation_compute_forces; src/simulation_checkrebuild.f90 simulation_checkrebuild;
! src/simulation_move_particles_vv.f90 simulation_move_particles_vv;
! src/simulation_move_particles_lfv.f90 simulation_move_particles_lfv; src/messenger.MPI.f90
! sendmols, messenger_updateborders) when building the reference with FLOWMOL_HIP.  The names,
! argument lists and abort behaviour are those of the reference; the state stays on the GPU
! and the host arrays r, v, a are refreshed where the reference's other consumers read them.
!
! This file uses the reference's modules (arrays_MD, physical_constants_MD,
! computational_constants_MD) and is compiled inside the reference's own build.
!-----------------------------------------------------------------------------------------------
module flowmol_hip_driver
    use iso_c_binding
    use flowmol_hip_c
    implicit none
    type(c_ptr), save :: fm_ctx = c_null_ptr
contains
    subroutine fm_check(rc)
        use interfaces, only : error_abort
        integer(c_int), intent(in) :: rc
        if (rc /= 0) call error_abort(fm_error_message())
    end subroutine fm_check

    ! called once after setup_set_parameters (src/md.f90:58): parameters and initial state
    subroutine fm_setup()
        use physical_constants_MD, only : np, density, wallslidev
        use computational_constants_MD, only : globaldomain, rcutoff, delta_rneighbr, delta_t, tplot, &
                                               rebuild_criteria, fixed_rebuild, potential_flag, &
                                               Mie_potential, vflux_outflag, pressure_outflag, &
                                               integration_algorithm, ensemble, tag_move, &
                                               teth_k2, teth_k4, teth_k6, binspercell, cv_conserve, &
                                               sort_flag, sort_freq, sortblocksize
        use arrays_MD, only : r, v, a, tag, rtether
        type(fm_params) :: p
        p%domain = globaldomain
        p%density = density
        p%rcutoff = rcutoff
        p%delta_rneighbr = delta_rneighbr
        p%delta_t = delta_t
        p%tplot = tplot
        p%teval = tplot
        p%rebuild_criteria = rebuild_criteria
        p%fixed_rebuild = fixed_rebuild
        p%potential_flag = potential_flag
        p%mie_potential = Mie_potential
        p%ntypes = 1
        p%mass_lookup = 1.d0
        p%vflux_outflag = vflux_outflag
        p%pressure_outflag = pressure_outflag
        p%integration_algorithm = integration_algorithm
        p%ensemble = ensemble
        p%teth_k2 = teth_k2
        p%teth_k4 = teth_k4
        p%teth_k6 = teth_k6
        p%wallslidev = wallslidev
        p%binspercell = binspercell
        p%cv_conserve = cv_conserve
        p%sort_flag = sort_flag
        p%sort_freq = sort_freq
        p%sortblocksize = sortblocksize
        call fm_check(fm_init(p, int(np, c_int32_t), int(np, c_int32_t), fm_ctx))
        call fm_check(fm_upload_state(fm_ctx, int(np, c_int32_t), r, v, a))
        if (ensemble == tag_move) call fm_check(fm_upload_tags(fm_ctx, tag, rtether))
    end subroutine fm_setup
end module flowmol_hip_driver

subroutine assign_to_cell
    use flowmol_hip_driver
    call fm_check(fm_assign_to_cell(fm_ctx))
end subroutine assign_to_cell

subroutine assign_to_neighbourlist
    use flowmol_hip_driver
    call fm_check(fm_assign_to_neighbourlist(fm_ctx))
end subroutine assign_to_neighbourlist

subroutine linklist_deallocateall
    use flowmol_hip_driver
    call fm_check(fm_linklist_deallocateall(fm_ctx))
end subroutine linklist_deallocateall

subroutine sort_mols
    use flowmol_hip_driver
    call fm_check(fm_sort_mols(fm_ctx))
end subroutine sort_mols

subroutine sendmols
    use flowmol_hip_driver
    call fm_check(fm_sendmols(fm_ctx))
end subroutine sendmols

subroutine messenger_updateborders(rebuild)
    use flowmol_hip_driver
    integer :: rebuild
    call fm_check(fm_updateborders(fm_ctx, int(rebuild, c_int32_t)))
end subroutine messenger_updateborders

subroutine simulation_checkrebuild(rebuild)
    use flowmol_hip_driver
    use computational_constants_MD, only : iter
    integer, intent(out) :: rebuild
    integer(c_int32_t) :: rb
    call fm_check(fm_checkrebuild(fm_ctx, int(iter, c_int32_t), rb))
    rebuild = rb
end subroutine simulation_checkrebuild

subroutine simulation_move_particles_vv(pass_num)
    use flowmol_hip_driver
    integer, intent(in) :: pass_num
    call fm_check(fm_move_particles_vv(fm_ctx, int(pass_num, c_int32_t)))
end subroutine simulation_move_particles_vv

subroutine simulation_move_particles_lfv
    use flowmol_hip_driver
    call fm_check(fm_move_particles_lfv(fm_ctx))
end subroutine simulation_move_particles_lfv

subroutine simulation_compute_forces
    use flowmol_hip_driver
    use computational_constants_MD, only : iter, tplot
    use arrays_MD, only : r, v, a, potenergymol, virialmol
    call fm_check(fm_compute_forces(fm_ctx, int(iter, c_int32_t)))
    ! record steps: the host-side consumers (simulation_record) read a, potenergymol, virialmol
    if (mod(iter, tplot) == 0) then
        call fm_check(fm_download_state(fm_ctx, r, v, a))
        call fm_check(fm_download_molecular(fm_ctx, potenergymol, virialmol))
    end if
end subroutine simulation_compute_forces
