!-----------------------------------------------------------------------------------------------
! Drop-in bodies for the reference's force / neighbour / integration routines.
!
! Link these in place of the reference's own definitions (src/linklist.f90 assign_to_cell,
! assign_to_neighbourlist, linklist_deallocateall, sort_mols; src/simulation_compute_forces.f90
! simulation_compute_forces; src/simulation_checkrebuild.f90 simulation_checkrebuild;
! src/simulation_move_particles_vv.f90 simulation_move_particles_vv;
! src/simulation_move_particles_lfv.f90 simulation_move_particles_lfv; src/messenger.MPI.f90
! sendmols, messenger_updateborders).  The names, argument lists and abort behaviour are those
! of the reference and src/md.f90 stays unchanged:
!
!  * the device context is created at the first call of any of these routines, from the host
!    arrays and the parameters setup_set_parameters has computed (fm_setup);
!  * until the time loop starts (the first call of a move routine) the host arrays are the
!    current state: assign_to_cell uploads them before binning, so setup routines that move
!    molecules and bin them again keep working;
!  * from then on the device holds the state.  The host arrays r, v, a, the per-molecule tallies
!    and the planes are refreshed where simulation_record reads them (record steps: after
!    simulation_move_particles_vv(2) for velocity Verlet, src/md.f90:159-161; after the force call
!    for leap-frog, src/md.f90:114-115), and after every rebuild that does not come from
!    simulation_checkrebuild (the pre-write rebuild of parallel_io_final_state,
!    src/parallel_io.MPI2.f90:1517-1523, which then writes the host arrays);
!  * simulation_checkrebuild keeps the reference's rescue snapshot on a CPU-time cadence and its
!    ABORTABORT file check (src/simulation_checkrebuild.f90:36-81);
!  * host consumers of the reference's linked lists (cell%head, neighbour%head) that the device
!    path does not fill are refused at setup with a clear message instead of reading empty lists.
!
! This file uses the reference's modules and is compiled inside the reference's own build.
!-----------------------------------------------------------------------------------------------
module flowmol_hip_driver
    use iso_c_binding
    use flowmol_hip_c
    implicit none
    type(c_ptr), save :: fm_ctx = c_null_ptr
    logical, save :: dev_owns = .false.      ! the time loop has started: the device state is current
    logical, save :: loop_rebuild = .false.  ! a rebuild requested by simulation_checkrebuild is running
    integer, save :: ntypes_dev = 0
    integer, save :: type_of(64) = -1        ! reference moltype -> device type (Mie potential)
    integer, save :: moltype_of(FM_MAX_TYPES) = 0
contains
    subroutine fm_check(rc)
        use interfaces, only : error_abort
        integer(c_int), intent(in) :: rc
        if (rc /= 0) call error_abort(fm_error_message())
    end subroutine fm_check

    subroutine fm_ensure()
        if (.not. c_associated(fm_ctx)) call fm_setup()
    end subroutine fm_ensure

    ! Host readers of the linked lists that the device path leaves empty: refuse them.
    subroutine fm_refuse_list_consumers()
        use interfaces, only : error_abort
        use computational_constants_MD, only : force_list, ensemble, nve, tag_move, cluster_analysis_outflag, &
                                               potential_flag, Mie_potential, pressure_outflag, CVforce_flag, VOID, &
                                               integration_algorithm, leap_frog_verlet, velocity_verlet
        use polymer_info_MD, only : solvent_flag, angular_potential
        if (force_list /= 3) call error_abort("FLOWMOL_HIP: only FORCE_LIST 3 (half-interaction neighbour lists) " // &
                                              "runs on the device")
        if (ensemble /= nve .and. ensemble /= tag_move) &
            call error_abort("FLOWMOL_HIP: thermostatted ensembles read the host neighbour lists; use NVE or tag_move")
        if (ensemble == tag_move .and. integration_algorithm /= leap_frog_verlet) &
            call error_abort("FLOWMOL_HIP: tag_move needs the leap-frog integrator, as in the reference")
        if (integration_algorithm /= leap_frog_verlet .and. integration_algorithm /= velocity_verlet) &
            call error_abort("FLOWMOL_HIP: unknown integration algorithm")
        if (cluster_analysis_outflag /= 0) &
            call error_abort("FLOWMOL_HIP: cluster analysis reads the host cell and neighbour lists")
        if (CVforce_flag /= VOID) call error_abort("FLOWMOL_HIP: CV forces read the host cell lists")
        if (pressure_outflag > 1) call error_abort("FLOWMOL_HIP: pressure_outflag > 1 reads the host lists")
        if (potential_flag == 1) then
            if (solvent_flag /= 0) call error_abort("FLOWMOL_HIP: explicit solvent is not on the device path")
            if (angular_potential /= 0) call error_abort("FLOWMOL_HIP: angular potentials are not on the device path")
            if (Mie_potential == 1) &
                call error_abort("FLOWMOL_HIP: Mie mode uses harmonic bonds, which are not on the device path")
        end if
    end subroutine fm_refuse_list_consumers

    ! Parameters and initial state (every fm_params field is set here).
    subroutine fm_setup()
        use interfaces, only : error_abort
        use physical_constants_MD, only : np, density, rcutoff, wallslidev
        use computational_constants_MD, only : globaldomain, delta_rneighbr, delta_t, tplot, teval, &
                                               rebuild_criteria, fixed_rebuild, potential_flag, Mie_potential, &
                                               vflux_outflag, pressure_outflag, integration_algorithm, ensemble, &
                                               tag_move, teth_k2, teth_k4, teth_k6, binspercell, cv_conserve, &
                                               sort_flag, sort_freq, sortblocksize
        use calculated_properties_MD, only : nplanes, planespacing
        use polymer_info_MD, only : k_c, R_0
        use module_set_parameters, only : mass_lookup, epsilon_lookup, lambdar_lookup, lambdaa_lookup, &
                                          C_lookup, alpha_lookup, potshift_lookup
        use arrays_MD, only : moltype
        type(fm_params) :: p
        integer :: n, a, b, k, t
        character(len=32) :: env
        integer :: envlen, envstat

        call fm_refuse_list_consumers()
        p%domain = globaldomain
        p%density = density
        p%rcutoff = rcutoff
        p%delta_rneighbr = delta_rneighbr
        p%delta_t = delta_t
        p%tplot = tplot
        p%teval = teval
        p%rebuild_criteria = rebuild_criteria
        p%fixed_rebuild = fixed_rebuild
        p%potential_flag = potential_flag
        p%mie_potential = Mie_potential
        ! FP32 pair law with FP64 state and accumulators when FLOWMOL_HIP_PRECISION=mixed
        p%precision = FM_PREC_FP64
        call get_environment_variable("FLOWMOL_HIP_PRECISION", env, envlen, envstat)
        if (envstat == 0 .and. env(1:envlen) == "mixed") p%precision = FM_PREC_MIXED
        p%mass_lookup = 1.d0
        p%epsilon_lookup = 0.d0
        p%lambdar_lookup = 0.d0
        p%lambdaa_lookup = 0.d0
        p%C_lookup = 0.d0
        p%alpha_lookup = 0.d0
        p%potshift_lookup = 0.d0
        p%rcutoff2_lookup = rcutoff**2
        p%ntypes = 1
        if (Mie_potential == 1) then
            ! the types present, in order of first appearance, become device types 0..ntypes-1;
            ! the reference has one global cutoff (src/setup_set_parameters.f90:946)
            ntypes_dev = 0
            type_of = -1
            do n = 1, np
                t = moltype(n)
                if (t < 1 .or. t > size(type_of)) call error_abort("FLOWMOL_HIP: moltype out of range")
                if (type_of(t) < 0) then
                    if (ntypes_dev == FM_MAX_TYPES) &
                        call error_abort("FLOWMOL_HIP: more molecule types than the device tables hold")
                    type_of(t) = ntypes_dev
                    ntypes_dev = ntypes_dev + 1
                    moltype_of(ntypes_dev) = t
                end if
            end do
            p%ntypes = max(ntypes_dev, 1)
            do a = 1, ntypes_dev
                p%mass_lookup(a) = mass_lookup(moltype_of(a))
                do b = 1, ntypes_dev
                    k = (a - 1) * FM_MAX_TYPES + b
                    p%epsilon_lookup(k) = epsilon_lookup(moltype_of(a), moltype_of(b))
                    p%lambdar_lookup(k) = lambdar_lookup(moltype_of(a), moltype_of(b))
                    p%lambdaa_lookup(k) = lambdaa_lookup(moltype_of(a), moltype_of(b))
                    p%C_lookup(k) = C_lookup(moltype_of(a), moltype_of(b))
                    p%alpha_lookup(k) = alpha_lookup(moltype_of(a), moltype_of(b))
                    p%potshift_lookup(k) = potshift_lookup(moltype_of(a), moltype_of(b))
                end do
            end do
        end if
        p%k_c = k_c
        p%R_0 = R_0
        p%ensemble = ensemble
        p%integration_algorithm = integration_algorithm
        p%teth_k2 = teth_k2
        p%teth_k4 = teth_k4
        p%teth_k6 = teth_k6
        p%wallslidev = wallslidev
        p%vflux_outflag = vflux_outflag
        p%pressure_outflag = pressure_outflag
        p%nplanes = 0
        p%planespacing = 0.d0
        if (vflux_outflag >= 1 .and. vflux_outflag <= 3) then
            p%nplanes = nplanes
            p%planespacing = planespacing
        end if
        p%pad1 = 0
        p%max_neighbours = 0            ! the library's default, grown on demand
        p%max_cell_np = 0
        p%brick = 0                     ! the library's default brick shape
        p%pad2 = 0
        p%binspercell = binspercell
        p%cv_conserve = cv_conserve
        p%pad3 = 0
        p%sort_flag = sort_flag
        p%sort_freq = sort_freq
        p%sortblocksize = sortblocksize
        p%pad4 = 0
        if (ensemble /= tag_move) p%teth_k2 = 0.d0
        call fm_check(fm_init(p, int(np, c_int32_t), int(np, c_int32_t), fm_ctx))
        call fm_upload_host()
    end subroutine fm_setup

    ! host arrays -> device (setup phase, and before every binning until the time loop starts)
    subroutine fm_upload_host()
        use physical_constants_MD, only : np
        use computational_constants_MD, only : ensemble, tag_move, potential_flag, Mie_potential
        use arrays_MD, only : r, v, a, tag, rtether, moltype
        use polymer_info_MD, only : monomer
        integer(c_int32_t), allocatable :: ibuf(:), c1(:), c2(:), c3(:), c4(:), c5(:)
        integer :: n
        call fm_check(fm_upload_state(fm_ctx, int(np, c_int32_t), r, v, a))
        if (Mie_potential == 1) then
            allocate(ibuf(np))
            do n = 1, np
                ibuf(n) = type_of(moltype(n))
            end do
            call fm_check(fm_upload_moltype(fm_ctx, ibuf))
        end if
        if (ensemble == tag_move) call fm_check(fm_upload_tags(fm_ctx, tag, rtether))
        if (potential_flag == 1) then
            allocate(c1(np), c2(np), c3(np), c4(4 * np), c5(np))
            do n = 1, np
                c1(n) = monomer(n)%chainID
                c2(n) = monomer(n)%subchainID
                c3(n) = monomer(n)%funcy
                c4(4 * n - 3:4 * n) = monomer(n)%bin_bflag(1:4)
                c5(n) = monomer(n)%glob_no
            end do
            call fm_check(fm_upload_monomers(fm_ctx, c1, c2, c3, c4, c5))
        end if
    end subroutine fm_upload_host

    ! device -> host arrays: state and the per-molecule labels in the current numbering
    subroutine fm_download_host()
        use physical_constants_MD, only : np
        use computational_constants_MD, only : ensemble, tag_move, potential_flag, Mie_potential
        use arrays_MD, only : r, v, a, tag, rtether, moltype
        use polymer_info_MD, only : monomer
        integer(c_int32_t), allocatable, target :: t1(:), m1(:), c1(:), c2(:), c3(:), c4(:), c5(:)
        type(c_ptr) :: pt, pm
        integer :: n
        call fm_check(fm_download_state(fm_ctx, r, v, a))
        if (ensemble == tag_move) call fm_check(fm_download_tags(fm_ctx, rtether))
        if (ensemble == tag_move .or. Mie_potential == 1) then
            allocate(t1(np), m1(np))
            pt = c_null_ptr
            pm = c_null_ptr
            if (ensemble == tag_move) pt = c_loc(t1)
            if (Mie_potential == 1) pm = c_loc(m1)
            call fm_check(fm_download_labels(fm_ctx, pt, pm, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                             c_null_ptr))
            do n = 1, np
                if (ensemble == tag_move) tag(n) = t1(n)
                if (Mie_potential == 1) moltype(n) = moltype_of(m1(n) + 1)
            end do
        end if
        if (potential_flag == 1) then
            allocate(c1(np), c2(np), c3(np), c4(4 * np), c5(np))
            call fm_check(fm_download_labels(fm_ctx, c_null_ptr, c_null_ptr, c_loc(c1), c_loc(c2), c_loc(c3), &
                                             c_loc(c4), c_loc(c5)))
            do n = 1, np
                monomer(n)%chainID = c1(n)
                monomer(n)%subchainID = c2(n)
                monomer(n)%funcy = c3(n)
                monomer(n)%bin_bflag(1:4) = c4(4 * n - 3:4 * n)
                monomer(n)%glob_no = c5(n)
            end do
        end if
    end subroutine fm_download_host

    ! What simulation_record reads at a record step: the state, the per-molecule tallies, and
    ! the stress tallies accumulated on the device since the last refresh (added to the host's
    ! own arrays, which the reference's output routines average and zero).
    subroutine fm_download_record()
        use physical_constants_MD, only : np
        use computational_constants_MD, only : vflux_outflag, pressure_outflag, potential_flag
        use arrays_MD, only : potenergymol, potenergymol_LJ, potenergymol_POLY, virialmol
        use calculated_properties_MD, only : nplanes, Pxy_plane, rfmol, Pxyface
        real(c_double), allocatable :: buf(:), pe(:), pp(:)
        integer :: n, i, j
        call fm_download_host()
        allocate(pe(np), pp(np))
        call fm_check(fm_download_potenergy_parts(fm_ctx, pe, pp))
        potenergymol_LJ(1:np) = pe
        if (potential_flag == 1) potenergymol_POLY(1:np) = pp
        potenergymol(1:np) = pe + pp
        call fm_check(fm_download_molecular(fm_ctx, pe, virialmol))
        if (vflux_outflag >= 1 .and. vflux_outflag <= 3) then
            allocate(buf(3 * nplanes))
            call fm_check(fm_get_mop(fm_ctx, buf, int(nplanes, c_int32_t)))
            call fm_check(fm_reset_mop(fm_ctx))
            Pxy_plane(:, 1:nplanes) = Pxy_plane(:, 1:nplanes) + reshape(buf, [3, nplanes])
            deallocate(buf)
        else if (vflux_outflag == 4) then
            allocate(buf(size(Pxyface)))
            call fm_check(fm_get_cv_stress(fm_ctx, buf, int(size(Pxyface), c_int64_t)))
            call fm_check(fm_reset_cv_stress(fm_ctx))
            Pxyface = Pxyface + reshape(buf, shape(Pxyface))
            deallocate(buf)
        end if
        if (pressure_outflag == 1) then
            allocate(buf(9 * np))
            call fm_check(fm_download_rfmol(fm_ctx, buf))
            do n = 1, np
                do i = 1, 3
                    do j = 1, 3
                        rfmol(n, i, j) = buf(9 * (n - 1) + 3 * (i - 1) + j)
                    end do
                end do
            end do
        end if
    end subroutine fm_download_record
end module flowmol_hip_driver

subroutine assign_to_cell
    use flowmol_hip_driver
    call fm_ensure()
    if (.not. dev_owns) call fm_upload_host()
    call fm_check(fm_assign_to_cell(fm_ctx))
end subroutine assign_to_cell

subroutine assign_to_neighbourlist
    use flowmol_hip_driver
    call fm_ensure()
    call fm_check(fm_assign_to_neighbourlist(fm_ctx))
    ! a rebuild outside the time loop's own (the pre-write rebuild of parallel_io_final_state):
    ! the host arrays are read next
    if (dev_owns .and. .not. loop_rebuild) call fm_download_host()
    loop_rebuild = .false.
end subroutine assign_to_neighbourlist

subroutine linklist_deallocateall
    use flowmol_hip_driver
    call fm_ensure()
    call fm_check(fm_linklist_deallocateall(fm_ctx))
end subroutine linklist_deallocateall

subroutine sort_mols
    use flowmol_hip_driver
    call fm_ensure()
    call fm_check(fm_sort_mols(fm_ctx))
end subroutine sort_mols

subroutine sendmols
    use flowmol_hip_driver
    call fm_ensure()
    call fm_check(fm_sendmols(fm_ctx))
end subroutine sendmols

subroutine messenger_updateborders(rebuild)
    use flowmol_hip_driver
    integer :: rebuild
    call fm_ensure()
    call fm_check(fm_updateborders(fm_ctx, int(rebuild, c_int32_t)))
end subroutine messenger_updateborders

! src/simulation_checkrebuild.f90:17-174: the rescue snapshot and the ABORTABORT check run on
! the host as in the reference (their final_state write refreshes the host arrays through the
! drop-in rebuild sequence); the rebuild criteria run on the device.
subroutine simulation_checkrebuild(rebuild)
    use flowmol_hip_driver
    use interfaces, only : error_abort
    use computational_constants_MD, only : iter, prefix_dir, rescue_snapshot_freq, overwrite_rescue_snapshot, &
                                           irank, iroot
    integer, intent(out) :: rebuild
    integer(c_int32_t) :: rb
    logical, save :: just_written_snapshot = .true.
    real(kind(0.d0)), save :: t1
    real(kind(0.d0)) :: t2, dt
    logical :: abort
    call fm_ensure()
    dev_owns = .true.
    if (just_written_snapshot) then
        just_written_snapshot = .false.
        call cpu_time(t1)
    end if
    call cpu_time(t2)
    dt = t2 - t1
    if (dt > rescue_snapshot_freq) then
        if (irank == iroot) then
            print '(2(a,f10.5))', ' It`s been ', dt, ' seconds since last rescue and requested freq is ', &
                rescue_snapshot_freq
            if (overwrite_rescue_snapshot) then
                print '(a,i8)', ' Rescue microstate written to ./results/final_state at iter ', iter
            else
                print '(a,i8)', ' Rescue microstate written to interim_state at iter ', iter
            end if
        end if
        just_written_snapshot = .true.
        call parallel_io_final_state(.not. overwrite_rescue_snapshot)
    end if
    inquire(file=trim(prefix_dir)//'ABORTABORT', exist=abort)
    if (abort) then
        print *, 'File ABORTABORT detected. Writing final_state file...'
        call parallel_io_final_state(.false.)
        call error_abort('Restart file written. Simulation aborted.')
    end if
    call fm_check(fm_checkrebuild(fm_ctx, int(iter, c_int32_t), rb))
    rebuild = rb
    loop_rebuild = rb == 1
end subroutine simulation_checkrebuild

subroutine simulation_move_particles_vv(pass_num)
    use flowmol_hip_driver
    use computational_constants_MD, only : iter, tplot
    integer, intent(in) :: pass_num
    call fm_ensure()
    dev_owns = .true.
    call fm_check(fm_move_particles_vv(fm_ctx, int(pass_num, c_int32_t)))
    ! simulation_record follows pass 2 (src/md.f90:159-161): v(t+dt) on record steps
    if (pass_num == 2 .and. mod(iter, tplot) == 0) call fm_download_record()
end subroutine simulation_move_particles_vv

subroutine simulation_move_particles_lfv
    use flowmol_hip_driver
    call fm_ensure()
    dev_owns = .true.
    call fm_check(fm_move_particles_lfv(fm_ctx))
end subroutine simulation_move_particles_lfv

subroutine simulation_compute_forces
    use flowmol_hip_driver
    use computational_constants_MD, only : iter, tplot, integration_algorithm, leap_frog_verlet
    call fm_ensure()
    call fm_check(fm_compute_forces(fm_ctx, int(iter, c_int32_t)))
    ! the initial record of setup (src/setup_initial_record.f90:88) and, under leap-frog, the
    ! record that follows the force call (src/md.f90:114-115) read the host arrays
    if (mod(iter, tplot) == 0 .and. (.not. dev_owns .or. integration_algorithm == leap_frog_verlet)) &
        call fm_download_record()
end subroutine simulation_compute_forces
