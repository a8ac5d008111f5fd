!> iso_c_binding interfaces of libohpic.so (include/ohpic.h).
module m_ohpic_c
    use iso_c_binding
    use oh_type, only: oh_particle
    implicit none
    private
    public :: ohpic_config, ohpic_diag
    public :: ohpic_config_default, ohpic_create, ohpic_destroy, ohpic_last_error
    public :: ohpic_load_oh_particles, ohpic_store_oh_particles, ohpic_particle_count
    public :: ohpic_seed_uniform_maxwell, ohpic_deposit, ohpic_field_solve
    public :: ohpic_gather_push, ohpic_boundary, ohpic_step, ohpic_synchronize
    public :: ohpic_get_field, ohpic_put_field, ohpic_fft3d_r2r, ohpic_diagnostics
    public :: ohpic_energy
    public :: ohpic_cell_counts, ohpic_boris_batch, ohpic_interp_batch
    public :: ohpic_scatter_batch, ohpic_wrap_batch
    public :: ohpic_snapshot_open, ohpic_snapshot_write, ohpic_snapshot_write_host
    public :: ohpic_snapshot_close

    integer, parameter, public :: OHPIC_MAX_SPECIES = 8
    integer, parameter, public :: OHPIC_FP64 = 0, OHPIC_MIXED = 1
    integer, parameter, public :: OHPIC_WEIGHTS_REFERENCE = 0, OHPIC_WEIGHTS_STANDARD = 1
    integer, parameter, public :: OHPIC_FIELD_RHO = 0, OHPIC_FIELD_PHI = 1, OHPIC_FIELD_EX = 2, &
                                  OHPIC_FIELD_EY = 3, OHPIC_FIELD_EZ = 4, OHPIC_FIELD_EB = 5

    type, bind(c) :: ohpic_config
        integer(c_int32_t) :: nx, ny, nz
        integer(c_int32_t) :: nspecies
        real(c_double) :: qm(OHPIC_MAX_SPECIES)
        real(c_double) :: q_macro(OHPIC_MAX_SPECIES)
        integer(c_int64_t) :: capacity(OHPIC_MAX_SPECIES)
        real(c_double) :: dt
        real(c_double) :: b0(3)
        integer(c_int32_t) :: precision
        integer(c_int32_t) :: weight_order
        integer(c_int32_t) :: tile(3)
        real(c_double) :: slack
        integer(c_int32_t) :: chunks
        integer(c_int32_t) :: track_ids
        integer(c_int32_t) :: device
        integer(c_int32_t) :: charge_bits
    end type

    type, bind(c) :: ohpic_diag
        integer(c_int64_t) :: step
        real(c_double) :: ke_old(OHPIC_MAX_SPECIES)
        real(c_double) :: ke_new(OHPIC_MAX_SPECIES)
        real(c_double) :: field_energy
        real(c_double) :: field_energy_e2
        real(c_double) :: net_charge
        real(c_double) :: total_energy
        integer(c_int64_t) :: repacks
        integer(c_int64_t) :: fast_particles
        integer(c_int64_t) :: rebalances
        integer(c_int64_t) :: compactions
    end type

    interface
        integer(c_int) function ohpic_config_default(cfg) bind(c, name='ohpic_config_default')
            import :: c_int, ohpic_config
            type(ohpic_config), intent(out) :: cfg
        end function

        integer(c_int) function ohpic_create(cfg, ctx) bind(c, name='ohpic_create')
            import :: c_int, c_ptr, ohpic_config
            type(ohpic_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function ohpic_destroy(ctx) bind(c, name='ohpic_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        type(c_ptr) function ohpic_last_error(ctx) bind(c, name='ohpic_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function ohpic_load_oh_particles(ctx, n, rec) &
            bind(c, name='ohpic_load_oh_particles')
            import :: c_int, c_ptr, c_int64_t, oh_particle
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            type(oh_particle), intent(in) :: rec(*)
        end function

        integer(c_int) function ohpic_store_oh_particles(ctx, cap, n, rec) &
            bind(c, name='ohpic_store_oh_particles')
            import :: c_int, c_ptr, c_int64_t, oh_particle
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: cap
            integer(c_int64_t), intent(out) :: n
            type(oh_particle), intent(inout) :: rec(*)
        end function

        integer(c_int) function ohpic_particle_count(ctx, ispec, n) &
            bind(c, name='ohpic_particle_count')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: ispec
            integer(c_int64_t), intent(out) :: n
        end function

        integer(c_int) function ohpic_seed_uniform_maxwell(ctx, initstate, initseq, np, vth) &
            bind(c, name='ohpic_seed_uniform_maxwell')
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: initstate, initseq
            integer(c_int64_t), intent(in) :: np(*)
            real(c_double), intent(in) :: vth(*)
        end function

        integer(c_int) function ohpic_deposit(ctx) bind(c, name='ohpic_deposit')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function ohpic_field_solve(ctx) bind(c, name='ohpic_field_solve')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function ohpic_gather_push(ctx) bind(c, name='ohpic_gather_push')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function ohpic_boundary(ctx) bind(c, name='ohpic_boundary')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function ohpic_step(ctx, nsteps) bind(c, name='ohpic_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
        end function

        integer(c_int) function ohpic_synchronize(ctx) bind(c, name='ohpic_synchronize')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function ohpic_get_field(ctx, which, values, nelem, lb, ub, comp) &
            bind(c, name='ohpic_get_field')
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: which, nelem, comp
            real(c_double), intent(inout) :: values(*)
            integer(c_int32_t), intent(in) :: lb(3), ub(3)
        end function

        integer(c_int) function ohpic_put_field(ctx, which, values, nelem, lb, ub, comp) &
            bind(c, name='ohpic_put_field')
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: which, nelem, comp
            real(c_double), intent(in) :: values(*)
            integer(c_int32_t), intent(in) :: lb(3), ub(3)
        end function

        integer(c_int) function ohpic_fft3d_r2r(device, nx, ny, nz, backward, in, out) &
            bind(c, name='ohpic_fft3d_r2r')
            import :: c_int, c_double
            integer(c_int), value :: device, nx, ny, nz, backward
            real(c_double), intent(in) :: in(*)
            real(c_double), intent(inout) :: out(*)
        end function

        integer(c_int) function ohpic_diagnostics(ctx, diag) bind(c, name='ohpic_diagnostics')
            import :: c_int, c_ptr, ohpic_diag
            type(c_ptr), value :: ctx
            type(ohpic_diag), intent(out) :: diag
        end function

        ! [kinetic, field, total, net charge] of the last step
        integer(c_int) function ohpic_energy(ctx, out) bind(c, name='ohpic_energy')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(4)
        end function

        integer(c_int) function ohpic_cell_counts(ctx, ispec, counts) &
            bind(c, name='ohpic_cell_counts')
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: ispec
            integer(c_int32_t), intent(inout) :: counts(*)
        end function

        integer(c_int) function ohpic_boris_batch(device, n, x, y, z, vx, vy, vz, eb, qm, dt) &
            bind(c, name='ohpic_boris_batch')
            import :: c_int, c_int64_t, c_double
            integer(c_int), value :: device
            integer(c_int64_t), value :: n
            real(c_double), intent(inout) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
            real(c_double), intent(in) :: eb(*)
            real(c_double), value :: qm, dt
        end function

        integer(c_int) function ohpic_interp_batch(device, n, pos, values, nelem, lb, ub, &
                                                   weight_order, out) &
            bind(c, name='ohpic_interp_batch')
            import :: c_int, c_int64_t, c_double, c_int32_t
            integer(c_int), value :: device
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: pos(*), values(*)
            integer(c_int), value :: nelem, weight_order
            integer(c_int32_t), intent(in) :: lb(3), ub(3)
            real(c_double), intent(inout) :: out(*)
        end function

        integer(c_int) function ohpic_scatter_batch(device, n, pos, values, nelem, lb, ub, &
                                                    amount, weight_order) &
            bind(c, name='ohpic_scatter_batch')
            import :: c_int, c_int64_t, c_double, c_int32_t
            integer(c_int), value :: device
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: pos(*), amount(*)
            real(c_double), intent(inout) :: values(*)
            integer(c_int), value :: nelem, weight_order
            integer(c_int32_t), intent(in) :: lb(3), ub(3)
        end function

        integer(c_int) function ohpic_wrap_batch(device, n, p, nn, applications) &
            bind(c, name='ohpic_wrap_batch')
            import :: c_int, c_int64_t, c_double, c_int32_t
            integer(c_int), value :: device
            integer(c_int64_t), value :: n
            real(c_double), intent(inout) :: p(*)
            integer(c_int32_t), value :: nn
            integer(c_int), value :: applications
        end function

        integer(c_int) function ohpic_snapshot_open(filename, group, global_shape, &
                                                    global_offset, out) &
            bind(c, name='ohpic_snapshot_open')
            import :: c_int, c_char, c_int64_t, c_ptr
            character(kind=c_char), intent(in) :: filename(*), group(*)
            integer(c_int64_t), intent(in) :: global_shape(3), global_offset(3)
            type(c_ptr), intent(out) :: out
        end function

        integer(c_int) function ohpic_snapshot_write(snap, dataset, ctx, which) &
            bind(c, name='ohpic_snapshot_write')
            import :: c_int, c_char, c_ptr
            type(c_ptr), value :: snap, ctx
            character(kind=c_char), intent(in) :: dataset(*)
            integer(c_int), value :: which
        end function

        integer(c_int) function ohpic_snapshot_write_host(snap, dataset, values, local_shape, &
                                                          local_offset) &
            bind(c, name='ohpic_snapshot_write_host')
            import :: c_int, c_char, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: snap
            character(kind=c_char), intent(in) :: dataset(*)
            real(c_double), intent(in) :: values(*)
            integer(c_int64_t), intent(in) :: local_shape(3), local_offset(3)
        end function

        integer(c_int) function ohpic_snapshot_close(snap) bind(c, name='ohpic_snapshot_close')
            import :: c_int, c_ptr
            type(c_ptr), value :: snap
        end function
    end interface
end module
