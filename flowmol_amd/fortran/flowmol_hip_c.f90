!-----------------------------------------------------------------------------------------------
! flowmol_hip_c -- iso_c_binding interface to libflowmol_hip (include/flowmol_hip.h).
!
! The derived type fm_params mirrors the C struct field by field (bind(C) gives the C layout).
! Every routine returns 0 on success; on failure fm_last_error() holds the message that the
! reference would pass to error_abort.
!-----------------------------------------------------------------------------------------------
module flowmol_hip_c
    use iso_c_binding
    implicit none

    integer, parameter :: FM_MAX_TYPES = 4, FM_MAX_FUNCY = 4
    integer(c_int32_t), parameter :: FM_REBUILD_VMAX = 0, FM_REBUILD_FIXED = 1, FM_REBUILD_DISPLACEMENT = 2
    integer(c_int32_t), parameter :: FM_ENSEMBLE_NVE = 0, FM_ENSEMBLE_TAG_MOVE = 6
    integer(c_int32_t), parameter :: FM_INTEGRATION_LFV = 0, FM_INTEGRATION_VV = 1
    integer(c_int32_t), parameter :: FM_PREC_FP64 = 0, FM_PREC_MIXED = 1
    integer(c_int), parameter :: FM_EABORT = 8

    type, bind(C) :: fm_params
        real(c_double)     :: domain(3)
        real(c_double)     :: density
        real(c_double)     :: rcutoff
        real(c_double)     :: delta_rneighbr(3)
        real(c_double)     :: delta_t
        integer(c_int32_t) :: tplot, teval, rebuild_criteria, fixed_rebuild
        integer(c_int32_t) :: potential_flag, mie_potential, ntypes, precision
        real(c_double)     :: mass_lookup(FM_MAX_TYPES)
        real(c_double)     :: epsilon_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: lambdar_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: lambdaa_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: C_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: alpha_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: potshift_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: rcutoff2_lookup(FM_MAX_TYPES*FM_MAX_TYPES)
        real(c_double)     :: k_c, R_0
        integer(c_int32_t) :: ensemble, integration_algorithm
        real(c_double)     :: teth_k2, teth_k4, teth_k6
        real(c_double)     :: wallslidev(3)
        integer(c_int32_t) :: vflux_outflag, pressure_outflag, nplanes, pad1
        real(c_double)     :: planespacing
        integer(c_int32_t) :: max_neighbours, max_cell_np
        integer(c_int32_t) :: brick(3), pad2
        real(c_double)     :: binspercell(3)
        integer(c_int32_t) :: cv_conserve, pad3
        integer(c_int32_t) :: sort_flag, sort_freq, sortblocksize, pad4
    end type fm_params

    type, bind(C) :: fm_macros
        real(c_double) :: potenergysum, potenergysum_LJ, potenergysum_POLY, virial, mv2sum, vsum, msum
        real(c_double) :: kinenergy, potenergy, totenergy, temperature, pressure, max_potenergymol
    end type fm_macros

    interface
        integer(c_int) function fm_init(p, np, capacity, ctx) bind(C, name="fm_init")
            import :: c_int, c_int32_t, c_ptr, fm_params
            type(fm_params), intent(in)       :: p
            integer(c_int32_t), value         :: np, capacity
            type(c_ptr), intent(out)          :: ctx
        end function
        subroutine fm_finalize(ctx) bind(C, name="fm_finalize")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        type(c_ptr) function fm_last_error() bind(C, name="fm_last_error")
            import :: c_ptr
        end function
        integer(c_int) function fm_upload_state(ctx, np, r, v, a) bind(C, name="fm_upload_state")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value        :: ctx
            integer(c_int32_t), value :: np
            real(c_double), intent(in) :: r(*), v(*), a(*)
        end function
        integer(c_int) function fm_download_state(ctx, r, v, a) bind(C, name="fm_download_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: r(*), v(*), a(*)
        end function
        integer(c_int) function fm_download_molecular(ctx, pe, vir) bind(C, name="fm_download_molecular")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: pe(*), vir(*)
        end function
        integer(c_int) function fm_move_particles_vv(ctx, pass_num) bind(C, name="fm_move_particles_vv")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value        :: ctx
            integer(c_int32_t), value :: pass_num
        end function
        integer(c_int) function fm_move_particles_lfv(ctx) bind(C, name="fm_move_particles_lfv")
            import :: c_int, c_ptr
            type(c_ptr), value        :: ctx
        end function
        integer(c_int) function fm_upload_tags(ctx, tag, rtether) bind(C, name="fm_upload_tags")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value        :: ctx
            integer(c_int32_t), intent(in) :: tag(*)
            real(c_double), intent(in) :: rtether(*)
        end function
        integer(c_int) function fm_checkrebuild(ctx, iter, rebuild) bind(C, name="fm_checkrebuild")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value              :: ctx
            integer(c_int32_t), value       :: iter
            integer(c_int32_t), intent(out) :: rebuild
        end function
        integer(c_int) function fm_linklist_deallocateall(ctx) bind(C, name="fm_linklist_deallocateall")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_sendmols(ctx) bind(C, name="fm_sendmols")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_sort_mols(ctx) bind(C, name="fm_sort_mols")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_assign_to_cell(ctx) bind(C, name="fm_assign_to_cell")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_updateborders(ctx, rebuild) bind(C, name="fm_updateborders")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value        :: ctx
            integer(c_int32_t), value :: rebuild
        end function
        integer(c_int) function fm_assign_to_neighbourlist(ctx) bind(C, name="fm_assign_to_neighbourlist")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_compute_forces(ctx, iter) bind(C, name="fm_compute_forces")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value        :: ctx
            integer(c_int32_t), value :: iter
        end function
        integer(c_int) function fm_evaluate_macroscopic_properties(ctx, m) &
                bind(C, name="fm_evaluate_macroscopic_properties")
            import :: c_int, c_ptr, fm_macros
            type(c_ptr), value            :: ctx
            type(fm_macros), intent(out)  :: m
        end function
        integer(c_int) function fm_upload_moltype(ctx, moltype) bind(C, name="fm_upload_moltype")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value             :: ctx
            integer(c_int32_t), intent(in) :: moltype(*)
        end function
        integer(c_int) function fm_upload_monomers(ctx, chainID, subchainID, funcy, bin_bflag, glob_no) &
                bind(C, name="fm_upload_monomers")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value             :: ctx
            integer(c_int32_t), intent(in) :: chainID(*), subchainID(*), funcy(*), bin_bflag(*), glob_no(*)
        end function
        integer(c_int) function fm_download_tags(ctx, rtether) bind(C, name="fm_download_tags")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: rtether(*)
        end function
        ! any argument may be c_null_ptr
        integer(c_int) function fm_download_labels(ctx, tag, moltype, chainID, subchainID, funcy, bin_bflag, &
                glob_no) bind(C, name="fm_download_labels")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, tag, moltype, chainID, subchainID, funcy, bin_bflag, glob_no
        end function
        integer(c_int) function fm_download_potenergy_parts(ctx, pe_lj, pe_poly) &
                bind(C, name="fm_download_potenergy_parts")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: pe_lj(*), pe_poly(*)
        end function
        integer(c_int) function fm_download_rfmol(ctx, rfmol) bind(C, name="fm_download_rfmol")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: rfmol(*)
        end function
        integer(c_int) function fm_get_mop(ctx, pxy, nplanes) bind(C, name="fm_get_mop")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: pxy(*)
            integer(c_int32_t), value   :: nplanes
        end function
        integer(c_int) function fm_reset_mop(ctx) bind(C, name="fm_reset_mop")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_get_cv_stress(ctx, pxyface, n) bind(C, name="fm_get_cv_stress")
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value          :: ctx
            real(c_double), intent(out) :: pxyface(*)
            integer(c_int64_t), value   :: n
        end function
        integer(c_int) function fm_reset_cv_stress(ctx) bind(C, name="fm_reset_cv_stress")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fm_synchronize(ctx) bind(C, name="fm_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
    end interface

contains

    ! message of the last failed call as a Fortran string
    function fm_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: n
        call c_f_pointer(fm_last_error(), s, [1024])
        n = 0
        do while (n < 1024)
            if (s(n+1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function fm_error_message

end module flowmol_hip_c
