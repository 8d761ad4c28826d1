!> Drives libicar_adv.so from Fortran through the iso_c_binding module the
!! way ICAR's advection module would: domain arrays with ICAR's bounds
!! (ims:ime, kms:kme, jms:jme; u to ime+1, v to jme+1), bound with c_loc,
!! uploaded once, advanced, downloaded.  Input and output are raw streams:
!!   in : 20 int32 header (nx ny nz ims kms jms nspec order fct advect_density
!!        nsteps bc_x bc_y variant halo adv h_order v_order flux_corr rk3),
!!        dx dt (real32), then u v w density
!!        jacobian jacobian_u jacobian_v jacobian_w advection_dz and the
!!        species, each in Fortran order
!!   out: the species after nsteps x (advect, ring conditions)
program test_icar_adv
    use iso_c_binding
    use icar_adv_c
    implicit none
    integer(c_int) :: hdr(20)
    real(c_float) :: dx, dt
    integer :: nx, ny, nz, ims, ime, kms, kme, jms, jme, nspec, s, rc, u_in, u_out
    real(c_float), allocatable, target :: u(:,:,:), v(:,:,:), w(:,:,:), density(:,:,:), jaco(:,:,:)
    real(c_float), allocatable, target :: jaco_u(:,:,:), jaco_v(:,:,:), jaco_w(:,:,:), adz(:,:,:)
    real(c_float), allocatable, target :: species(:,:,:,:)
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
    read(u_in) u, v, w, density, jaco, jaco_u, jaco_v, jaco_w, adz, species
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
    call check(icar_adv_run(h, dt, hdr(11), hdr(12), hdr(13)))
    call check(icar_adv_download(h, ICAR_ADV_MASK_ALL_SPECIES))
    call check(icar_adv_destroy(h))

    open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u_out) species
    close(u_out)

contains

    subroutine check(code)
        integer(c_int), intent(in) :: code
        if (code /= 0) then
            print *, "icar_adv error ", code, ": ", icar_adv_error_message(h)
            error stop 1
        end if
    end subroutine

end program test_icar_adv
