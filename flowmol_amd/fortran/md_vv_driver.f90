!-----------------------------------------------------------------------------------------------
! md_vv_driver -- the reference's velocity Verlet time loop (src/md.f90:143-163) in Fortran,
! calling libflowmol_hip through flowmol_hip_c.  It reads a state written by the reference's
! final_state writer layout (positions and velocities) through the command line arguments:
!   md_vv_driver <nunits> <nsteps>
! builds the FCC lattice of setup_initialise_lattice (src/setup_initialise_microstate.f90:244-304)
! with zero velocities plus a deterministic perturbation, runs nsteps, and prints the
! macroscopic properties every tplot steps as the reference's stdout line does.
!-----------------------------------------------------------------------------------------------
program md_vv_driver
    use iso_c_binding
    use flowmol_hip_c
    implicit none
    type(fm_params) :: p
    type(fm_macros) :: m
    type(c_ptr) :: ctx
    integer :: nunits, nsteps, np, n, ix, iy, iz, j, iter
    integer(c_int32_t) :: rebuild
    real(c_double) :: a_lat, L, c(3), rc(3), mv2
    real(c_double), allocatable :: r(:,:), v(:,:), acc(:,:)
    character(len=32) :: arg

    call get_command_argument(1, arg); read(arg, *) nunits
    call get_command_argument(2, arg); read(arg, *) nsteps
    np = 4*nunits**3
    L = nunits/((0.8d0/4.d0)**(1.d0/3))
    a_lat = L/nunits
    allocate(r(3,np), v(3,np), acc(3,np))
    n = 0
    do iz = 1, nunits
    c(3) = (iz - 0.75d0)*a_lat
    do ix = 1, nunits
    c(1) = (ix - 0.75d0)*a_lat
    do iy = 1, nunits
    c(2) = (iy - 0.75d0)*a_lat
        do j = 1, 4
            rc = c
            if (j == 2) then; rc(1) = c(1) + 0.5d0*a_lat; rc(3) = c(3) + 0.5d0*a_lat; end if
            if (j == 3) then; rc(2) = c(2) + 0.5d0*a_lat; rc(3) = c(3) + 0.5d0*a_lat; end if
            if (j == 4) then; rc(1) = c(1) + 0.5d0*a_lat; rc(2) = c(2) + 0.5d0*a_lat; end if
            n = n + 1
            r(:,n) = rc - 0.5d0*L
            ! deterministic velocities: a fixed pattern with zero net momentum per basis cell
            v(1,n) = merge(1.d0, -1.d0, mod(n,2) == 0)
            v(2,n) = merge(1.d0, -1.d0, mod(n/2,2) == 0)
            v(3,n) = merge(1.d0, -1.d0, mod(n/4,2) == 0)
        end do
    end do
    end do
    end do
    acc = 0.d0

    p%domain = L
    p%density = 0.8d0
    p%rcutoff = 2.5d0
    p%delta_rneighbr = 0.3d0
    p%delta_t = 0.005d0
    p%tplot = 10
    p%teval = 10
    p%rebuild_criteria = FM_REBUILD_DISPLACEMENT
    p%fixed_rebuild = 20
    p%potential_flag = 0
    p%mie_potential = 0
    p%ntypes = 1
    p%precision = FM_PREC_FP64
    p%mass_lookup = 1.d0
    p%ensemble = FM_ENSEMBLE_NVE
    p%integration_algorithm = FM_INTEGRATION_VV
    p%vflux_outflag = 0
    p%pressure_outflag = 0
    p%nplanes = 0
    p%planespacing = 0.d0
    p%max_neighbours = 0
    p%max_cell_np = 0
    p%brick = 0
    p%binspercell = 1.d0
    p%cv_conserve = 0
    p%sort_flag = 0
    p%sort_freq = 1
    p%sortblocksize = 1

    call check(fm_init(p, int(np, c_int32_t), int(np, c_int32_t), ctx))
    call check(fm_upload_state(ctx, int(np, c_int32_t), r, v, acc))
    ! setup_MD: assign_to_cell, messenger_updateborders(1), assign_to_neighbourlist,
    ! then the initial force call of setup_initial_record
    call check(fm_assign_to_cell(ctx))
    call check(fm_updateborders(ctx, 1_c_int32_t))
    call check(fm_assign_to_neighbourlist(ctx))
    call check(fm_compute_forces(ctx, 0_c_int32_t))
    do iter = 1, nsteps
        call check(fm_move_particles_vv(ctx, 1_c_int32_t))
        call check(fm_updateborders(ctx, 0_c_int32_t))
        call check(fm_checkrebuild(ctx, int(iter, c_int32_t), rebuild))
        if (rebuild == 1) then
            call check(fm_linklist_deallocateall(ctx))
            call check(fm_sendmols(ctx))
            call check(fm_sort_mols(ctx))
            call check(fm_assign_to_cell(ctx))
            call check(fm_updateborders(ctx, 1_c_int32_t))
            call check(fm_assign_to_neighbourlist(ctx))
        end if
        call check(fm_compute_forces(ctx, int(iter, c_int32_t)))
        call check(fm_move_particles_vv(ctx, 2_c_int32_t))
        if (mod(iter, p%tplot) == 0) then
            ! simulation_record follows pass 2: the host refresh of a record step takes v(t+dt),
            ! the velocities the device's sums use
            call check(fm_evaluate_macroscopic_properties(ctx, m))
            call check(fm_download_state(ctx, r, v, acc))
            mv2 = sum(v(:, 1:np)**2)
            write(*,'(i8,a,5(f19.15,a),2(es24.16,a))') iter, ';', m%kinenergy, ';', m%potenergy, ';', &
                m%totenergy, ';', m%temperature, ';', m%pressure, ';', mv2, ';', m%mv2sum, ';'
        end if
    end do
    call check(fm_download_state(ctx, r, v, acc))
    write(*,'(a,3es24.16)') 'final r(:,1) ', r(:,1)
    call fm_finalize(ctx)

contains
    subroutine check(rc)
        integer(c_int), intent(in) :: rc
        if (rc /= 0) then
            write(*,*) 'flowmol_hip error: ', fm_error_message()
            stop 1
        end if
    end subroutine check
end program md_vv_driver
