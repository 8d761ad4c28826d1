!> ----------------------------------------------------------------------------
!!  iso_c_binding interface to libicar_adv.so, the MI355X MPDATA advection
!!  library (C-ABI declared in include/icar_adv.h).
!!
!!  ICAR's advection module (src/physics/advection_driver.f90) calls
!!  icar_adv_create from adv_init, binds domain%u/v/w/density, the jacobians
!!  and advection_dz plus every domain%adv_vars entry, and forwards
!!  advect(domain, options, dt) to icar_adv_advect for adv=2.
!! ----------------------------------------------------------------------------
module icar_adv_c
    use iso_c_binding
    implicit none
    private

    public :: icar_adv_grid, icar_adv_opts
    public :: icar_adv_create, icar_adv_destroy, icar_adv_bind_field, icar_adv_bind_species
    public :: icar_adv_upload, icar_adv_download, icar_adv_advect, icar_adv_mass_budget
    public :: icar_adv_max_courant, icar_adv_last_error, icar_adv_sync, icar_adv_apply_bc
    public :: icar_adv_divergence, icar_adv_set_compensation
    public :: icar_adv_run, icar_adv_set_graph, icar_adv_error_message
    public :: icar_adv_balance_w, icar_adv_compute_dt, icar_adv_launch_config
    public :: icar_adv_set_relax_filter, icar_adv_set_forcing, icar_adv_apply_forcing
    public :: icar_adv_halo_bytes, icar_adv_halo_pack, icar_adv_halo_unpack, icar_adv_halo_exchange
    public :: icar_adv_enforce_limits, icar_adv_snapshot, icar_adv_snapshot_wait
    public :: ICAR_ADV_NORTH, ICAR_ADV_SOUTH, ICAR_ADV_EAST, ICAR_ADV_WEST
    public :: ICAR_ADV_FIELD_U, ICAR_ADV_FIELD_V, ICAR_ADV_FIELD_W, ICAR_ADV_FIELD_DENSITY
    public :: ICAR_ADV_FIELD_JACOBIAN, ICAR_ADV_FIELD_JACOBIAN_U, ICAR_ADV_FIELD_JACOBIAN_V
    public :: ICAR_ADV_FIELD_JACOBIAN_W, ICAR_ADV_FIELD_ADVECTION_DZ
    public :: ICAR_ADV_STORE_F32, ICAR_ADV_STORE_BF16, ICAR_ADV_STORE_F16
    public :: ICAR_ADV_STORE_BF16_SR, ICAR_ADV_STORE_F16_SR
    public :: ICAR_ADV_MASK_ALL_FIELDS, ICAR_ADV_MASK_ALL_SPECIES
    public :: ICAR_ADV_OK, ICAR_ADV_EINVAL, ICAR_ADV_ESHAPE, ICAR_ADV_EUNBOUND, ICAR_ADV_ERANGE
    public :: ICAR_ADV_ENOMEM, ICAR_ADV_ECOURANT, ICAR_ADV_STD, ICAR_ADV_MPDATA

    integer(c_int), parameter :: ICAR_ADV_FIELD_U = 0, ICAR_ADV_FIELD_V = 1, ICAR_ADV_FIELD_W = 2
    integer(c_int), parameter :: ICAR_ADV_FIELD_DENSITY = 3, ICAR_ADV_FIELD_JACOBIAN = 4
    integer(c_int), parameter :: ICAR_ADV_FIELD_JACOBIAN_U = 5, ICAR_ADV_FIELD_JACOBIAN_V = 6
    integer(c_int), parameter :: ICAR_ADV_FIELD_JACOBIAN_W = 7, ICAR_ADV_FIELD_ADVECTION_DZ = 8
    integer(c_int), parameter :: ICAR_ADV_STORE_F32 = 0, ICAR_ADV_STORE_BF16 = 1, ICAR_ADV_STORE_F16 = 2
    integer(c_int), parameter :: ICAR_ADV_STORE_BF16_SR = 3, ICAR_ADV_STORE_F16_SR = 4
    integer(c_int), parameter :: ICAR_ADV_OK = 0, ICAR_ADV_EINVAL = -1, ICAR_ADV_ESHAPE = -2
    integer(c_int), parameter :: ICAR_ADV_EUNBOUND = -3, ICAR_ADV_ERANGE = -4, ICAR_ADV_ENOMEM = -5
    integer(c_int), parameter :: ICAR_ADV_ECOURANT = -6
    integer(c_int), parameter :: ICAR_ADV_STD = 1, ICAR_ADV_MPDATA = 2  ! kADV_STD, kADV_MPDATA
    integer(c_int), parameter :: ICAR_ADV_NORTH = 0, ICAR_ADV_SOUTH = 1, ICAR_ADV_EAST = 2
    integer(c_int), parameter :: ICAR_ADV_WEST = 3
    integer(c_long_long), parameter :: ICAR_ADV_MASK_ALL_FIELDS = int(z'1FF', c_long_long)
    ! bits 16..63 (species 0..47); written as a negative two's-complement value
    integer(c_long_long), parameter :: ICAR_ADV_MASK_ALL_SPECIES = -65536_c_long_long

    !> bounds and spacing copied by mpdata_init from domain_t (adv_mpdata.f90:422-431,486)
    type, bind(C) :: icar_adv_grid
        integer(c_int) :: ims, ime, kms, kme, jms, jme
        integer(c_int) :: its, ite, jts, jte
        real(c_float)  :: dx
    end type

    !> namelist switches: mpdata_order, MPDATA_FCT, advect_density, debug, RK3
    !! (options_obj.f90:1568,1579-1584,1025,2198), the kernel variant, and
    !! for the standard scheme adv, h_order, v_order, flux_corr
    !! (options_obj.f90:470-490,1582-1584)
    type, bind(C) :: icar_adv_opts
        integer(c_int) :: mpdata_order = 2
        integer(c_int) :: mpdata_fct = 1
        integer(c_int) :: advect_density = 0
        integer(c_int) :: debug = 0
        integer(c_int) :: rk3 = 0
        integer(c_int) :: kernel_variant = 1
        integer(c_int) :: advection = 2
        integer(c_int) :: h_order = 1
        integer(c_int) :: v_order = 1
        integer(c_int) :: flux_corr = 0
    end type

    interface
        integer(c_int) function icar_adv_create(grid, opts, handle) bind(C, name="icar_adv_create")
            import :: c_int, c_ptr, icar_adv_grid, icar_adv_opts
            type(icar_adv_grid), intent(in) :: grid
            type(icar_adv_opts), intent(in) :: opts
            type(c_ptr), intent(out) :: handle
        end function

        integer(c_int) function icar_adv_destroy(handle) bind(C, name="icar_adv_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_int) function icar_adv_bind_field(handle, field_id, host, lb, ub) &
                bind(C, name="icar_adv_bind_field")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: field_id
            type(c_ptr), value :: host
            integer(c_int), intent(in) :: lb(3), ub(3)
        end function

        integer(c_int) function icar_adv_bind_species(handle, idx, host, storage, pow2_exp) &
                bind(C, name="icar_adv_bind_species")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: idx, storage, pow2_exp
            type(c_ptr), value :: host
        end function

        integer(c_int) function icar_adv_upload(handle, mask) bind(C, name="icar_adv_upload")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: handle
            integer(c_long_long), value :: mask
        end function

        integer(c_int) function icar_adv_download(handle, mask) bind(C, name="icar_adv_download")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: handle
            integer(c_long_long), value :: mask
        end function

        integer(c_int) function icar_adv_advect(handle, dt) bind(C, name="icar_adv_advect")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: handle
            real(c_float), value :: dt
        end function

        integer(c_int) function icar_adv_mass_budget(handle, per_species) &
                bind(C, name="icar_adv_mass_budget")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle
            real(c_double), intent(out) :: per_species(*)
        end function

        integer(c_int) function icar_adv_max_courant(handle, out) bind(C, name="icar_adv_max_courant")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle
            real(c_double), intent(out) :: out(3)
        end function

        ! the fused kernels' flux divergence with compensated sums (on /= 0)
        integer(c_int) function icar_adv_set_compensation(handle, on) &
                bind(C, name="icar_adv_set_compensation")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: on
        end function

        ! test_divergence (adv_mpdata.f90:445-475) of the last step: out(1) cells
        ! with |div| > 1e-3, out(2) the largest |div|
        integer(c_int) function icar_adv_divergence(handle, out) bind(C, name="icar_adv_divergence")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle
            real(c_double), intent(out) :: out(2)
        end function

        type(c_ptr) function icar_adv_last_error(handle) bind(C, name="icar_adv_last_error")
            import :: c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_int) function icar_adv_sync(handle) bind(C, name="icar_adv_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_int) function icar_adv_apply_bc(handle, bc_x, bc_y) bind(C, name="icar_adv_apply_bc")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: bc_x, bc_y
        end function

        integer(c_int) function icar_adv_run(handle, dt, nsteps, bc_x, bc_y) bind(C, name="icar_adv_run")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: handle
            real(c_float), value :: dt
            integer(c_int), value :: nsteps, bc_x, bc_y
        end function

        integer(c_int) function icar_adv_set_graph(handle, enable) bind(C, name="icar_adv_set_graph")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: enable
        end function

        !> balance_uvw (wind.f90:88-164) on the device winds
        integer(c_int) function icar_adv_balance_w(handle) bind(C, name="icar_adv_balance_w")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        !> compute_dt (time_step.f90:51-168, use_density=.false.) on the device winds
        integer(c_int) function icar_adv_compute_dt(handle, cfl, cfl_strictness, dz_levels, dt) &
                bind(C, name="icar_adv_compute_dt")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: handle
            real(c_float), value :: cfl
            integer(c_int), value :: cfl_strictness
            real(c_float), intent(in) :: dz_levels(*)
            real(c_float), intent(out) :: dt
        end function

        !> init_relax_filters (domain_obj.f90:3285-3390) for the boundary flags
        integer(c_int) function icar_adv_set_relax_filter(handle, west, east, north, south, halo) &
                bind(C, name="icar_adv_set_relax_filter")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: west, east, north, south, halo
        end function

        !> a species' forcing_hi%data_3d and dqdt_3d (update_delta_fields)
        integer(c_int) function icar_adv_set_forcing(handle, idx, forcing_hi, dqdt) &
                bind(C, name="icar_adv_set_forcing")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: handle
            integer(c_int), value :: idx
            real(c_float), intent(in) :: forcing_hi(*), dqdt(*)
        end function

        !> apply_forcing (domain_obj.f90:3493-3503) of the forced species
        integer(c_int) function icar_adv_apply_forcing(handle, dt) bind(C, name="icar_adv_apply_forcing")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle
            real(c_double), value :: dt
        end function

        !> halo exchange of the advected species (domain%halo_exchange,
        !! domain_obj.f90:573-650): a side's strip into device memory, device
        !! memory into the halo on a side, or one handle's strip straight into
        !! its neighbour's opposite halo on the same device
        integer(c_int) function icar_adv_halo_bytes(handle, side, nbytes) &
                bind(C, name="icar_adv_halo_bytes")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: handle
            integer(c_int), value :: side
            integer(c_long_long), intent(out) :: nbytes
        end function

        integer(c_int) function icar_adv_halo_pack(handle, side, dev_buf, nbytes) &
                bind(C, name="icar_adv_halo_pack")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: handle, dev_buf
            integer(c_int), value :: side
            integer(c_long_long), value :: nbytes
        end function

        integer(c_int) function icar_adv_halo_unpack(handle, side, dev_buf, nbytes) &
                bind(C, name="icar_adv_halo_unpack")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: handle, dev_buf
            integer(c_int), value :: side
            integer(c_long_long), value :: nbytes
        end function

        integer(c_int) function icar_adv_halo_exchange(from, side, to) &
                bind(C, name="icar_adv_halo_exchange")
            import :: c_int, c_ptr
            type(c_ptr), value :: from, to
            integer(c_int), value :: side
        end function

        !> domain%enforce_limits (domain_obj.f90:3092-3150) on the device species
        integer(c_int) function icar_adv_enforce_limits(handle) bind(C, name="icar_adv_enforce_limits")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        !> output/restart hand-off: start copying the masked fields and species
        !! out (overlaps the following steps), then complete into the bound arrays
        integer(c_int) function icar_adv_snapshot(handle, mask) bind(C, name="icar_adv_snapshot")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: handle
            integer(c_long_long), value :: mask
        end function

        integer(c_int) function icar_adv_snapshot_wait(handle) bind(C, name="icar_adv_snapshot_wait")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        integer(c_int) function icar_adv_launch_config(handle, corr_cfg, ksplit) &
                bind(C, name="icar_adv_launch_config")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), intent(out) :: corr_cfg, ksplit
        end function
    end interface

contains

    !> The library's message for the last failure on handle, as a Fortran
    !! string (for `error stop`, as ICAR stops on errors, time_step.f90:165).
    function icar_adv_error_message(handle) result(msg)
        type(c_ptr), intent(in) :: handle
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: chars(:)
        type(c_ptr) :: p
        integer :: n
        p = icar_adv_last_error(handle)
        if (.not. c_associated(p)) then
            msg = ""
            return
        end if
        call c_f_pointer(p, chars, [4096])
        n = 0
        do while (n < 4096)
            if (chars(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(chars(1:n), msg)
    end function

end module icar_adv_c
