!> Drives the entry points around the step from Fortran, in the order an ICAR
!! time step meets them: balance_uvw, compute_dt, the advection steps and
!! the Courant maxima of the last one, relaxation-zone forcing,
!! enforce_limits, the mass and halo diagnostics, and the output hand-off
!! through snapshot/snapshot_wait.
!! Input and output are raw streams:
!!   in : the header, dx, dt, fields and species of test_icar_adv, then
!!        cfl (real32), cfl_strictness, the four relax flags west east north
!!        south (int32), the forcing dt (real64), dz_levels(nz) (real32), and
!!        forcing_hi, dqdt of species 1 (Fortran order, mass grid)
!!   out: the snapshot's w and species; max Courant (3 real64); compute_dt's
!!        dt (real32); mass budget (nspec real64); halo bytes of the north,
!!        south, east and west strips (int64); corr_cfg, ksplit (int32)
program test_icar_adv_ext
    use iso_c_binding
    use icar_adv_c
    implicit none
    integer(c_int) :: hdr(20), strictness, flags(4), corr_cfg, ksplit
    real(c_float) :: dx, dt, cfl, dt_cfl
    real(c_double) :: dt_forcing, courant(3)
    real(c_double), allocatable :: budget(:)
    integer(c_long_long) :: nbytes(4)
    integer :: nx, ny, nz, ims, ime, kms, kme, jms, jme, nspec, s, rc, u_in, u_out
    real(c_float), allocatable, target :: u(:,:,:), v(:,:,:), w(:,:,:), density(:,:,:), jaco(:,:,:)
    real(c_float), allocatable, target :: jaco_u(:,:,:), jaco_v(:,:,:), jaco_w(:,:,:), adz(:,:,:)
    real(c_float), allocatable, target :: species(:,:,:,:), forcing_hi(:,:,:), dqdt(:,:,:)
    real(c_float), allocatable :: dz_levels(:)
    type(icar_adv_grid) :: grid
    type(icar_adv_opts) :: opts
    type(c_ptr) :: h
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u_in) hdr, dx, dt
    nx = hdr(1); ny = hdr(2); nz = hdr(3)
    ims = hdr(4); kms = hdr(5); jms = hdr(6); nspec = hdr(7)
    ime = ims + nx - 1; kme = kms + nz - 1; jme = jms + ny - 1
    allocate(u(ims:ime+1, kms:kme, jms:jme), jaco_u(ims:ime+1, kms:kme, jms:jme))
    allocate(v(ims:ime, kms:kme, jms:jme+1), jaco_v(ims:ime, kms:kme, jms:jme+1))
    allocate(w(ims:ime, kms:kme, jms:jme), density(ims:ime, kms:kme, jms:jme))
    allocate(jaco(ims:ime, kms:kme, jms:jme), jaco_w(ims:ime, kms:kme, jms:jme))
    allocate(adz(ims:ime, kms:kme, jms:jme), species(ims:ime, kms:kme, jms:jme, nspec))
    allocate(forcing_hi(ims:ime, kms:kme, jms:jme), dqdt(ims:ime, kms:kme, jms:jme))
    allocate(dz_levels(kms:kme), budget(nspec))
    read(u_in) u, v, w, density, jaco, jaco_u, jaco_v, jaco_w, adz, species
    read(u_in) cfl, strictness, flags, dt_forcing, dz_levels, forcing_hi, dqdt
    close(u_in)

    grid = icar_adv_grid(ims, ime, kms, kme, jms, jme, ims + hdr(15), ime - hdr(15), &
                         jms + hdr(15), jme - hdr(15), dx)
    opts%mpdata_order = hdr(8)
    opts%mpdata_fct = hdr(9)
    opts%advect_density = hdr(10)
    opts%kernel_variant = hdr(14)
    opts%advection = hdr(16)
    opts%h_order = hdr(17)
    opts%v_order = hdr(18)
    opts%flux_corr = hdr(19)
    opts%rk3 = hdr(20)
    rc = icar_adv_create(grid, opts, h)
    if (rc /= 0) error stop "icar_adv_create failed"
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_U, c_loc(u), lbound(u), ubound(u)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_V, c_loc(v), lbound(v), ubound(v)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_W, c_loc(w), lbound(w), ubound(w)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_DENSITY, c_loc(density), lbound(density), &
                                   ubound(density)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_JACOBIAN, c_loc(jaco), lbound(jaco), ubound(jaco)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_JACOBIAN_U, c_loc(jaco_u), lbound(jaco_u), &
                                   ubound(jaco_u)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_JACOBIAN_V, c_loc(jaco_v), lbound(jaco_v), &
                                   ubound(jaco_v)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_JACOBIAN_W, c_loc(jaco_w), lbound(jaco_w), &
                                   ubound(jaco_w)))
    call check(icar_adv_bind_field(h, ICAR_ADV_FIELD_ADVECTION_DZ, c_loc(adz), lbound(adz), ubound(adz)))
    do s = 1, nspec
        call check(icar_adv_bind_species(h, s - 1, c_loc(species(ims, kms, jms, s)), &
                                         ICAR_ADV_STORE_F32, 0))
    end do
    call check(icar_adv_upload(h, ior(ICAR_ADV_MASK_ALL_FIELDS, ICAR_ADV_MASK_ALL_SPECIES)))

    call check(icar_adv_balance_w(h))
    call check(icar_adv_compute_dt(h, cfl, strictness, dz_levels, dt_cfl))
    call check(icar_adv_run(h, dt, hdr(11), hdr(12), hdr(13)))
    call check(icar_adv_max_courant(h, courant))
    call check(icar_adv_set_relax_filter(h, flags(1), flags(2), flags(3), flags(4), hdr(15)))
    call check(icar_adv_set_forcing(h, 0, forcing_hi, dqdt))
    call check(icar_adv_apply_forcing(h, dt_forcing))
    call check(icar_adv_enforce_limits(h))
    call check(icar_adv_mass_budget(h, budget))
    call check(icar_adv_halo_bytes(h, ICAR_ADV_NORTH, nbytes(1)))
    call check(icar_adv_halo_bytes(h, ICAR_ADV_SOUTH, nbytes(2)))
    call check(icar_adv_halo_bytes(h, ICAR_ADV_EAST, nbytes(3)))
    call check(icar_adv_halo_bytes(h, ICAR_ADV_WEST, nbytes(4)))
    call check(icar_adv_launch_config(h, corr_cfg, ksplit))
    call check(icar_adv_snapshot(h, ior(ICAR_ADV_MASK_ALL_FIELDS, ICAR_ADV_MASK_ALL_SPECIES)))
    call check(icar_adv_snapshot_wait(h))
    call check(icar_adv_destroy(h))

    open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u_out) w, species, courant, dt_cfl, budget, nbytes, corr_cfg, ksplit
    close(u_out)

contains

    subroutine check(code)
        integer(c_int), intent(in) :: code
        if (code /= 0) then
            print *, "icar_adv error ", code, ": ", icar_adv_error_message(h)
            error stop 1
        end if
    end subroutine

end program test_icar_adv_ext
