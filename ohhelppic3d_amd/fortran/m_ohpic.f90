!> GPU (MI355X) implementations of ohhelppic3d's per-timestep interfaces.
!>
!> Each type extends the reference's abstract type and keeps its deferred,
!> per-particle binding (served by a batched GPU kernel through the C ABI);
!> the whole-array bindings (scatter_all, move_all, apply_all, solve) act on
!> every particle of a device-resident t_OhpicSimulation at once and are what
!> a GPU mainstep calls instead of the reference's per-particle loops
!> (ohhelppic3d.f90:311-358).
!>
!>   reference                                  here
!>   t_ParticleMoverBoris   (particle_mover_boris.f90)   t_ParticleMoverHip
!>   t_LinearInterpolator   (linear_interpolator.f90)    t_InterpolatorHip
!>   t_LinearScatter        (linear_scatter.f90)         t_ScatterHip
!>   t_ParticleBoundaries   (particle_boundaries.f90)    t_ParticleBoundariesHip
!>   t_PoissonFieldSolver   (poisson_field_solver.f90)   t_PoissonFieldSolverHip
!>   t_MPIFFTW3Solver3d     (mpi_fftw3_solver.f90)       t_FFTSolverHip
!>   t_Hdf5ForOhfield       (hdf5_for_ohfield.f90)       t_Hdf5ForOhfieldHip
!>
!> Non-zero statuses end the program with `error stop` and the library's
!> message, as the reference's own failures do.
module m_ohpic
    use iso_c_binding
    use m_ohpic_c
    use oh_type, only: oh_particle
    use m_block, only: t_Block, new_Block
    use m_ohfield, only: t_OhField
    use m_ohhelp, only: t_OhHelp
    use m_particle_mover, only: t_ParticleMover
    use m_interpolator, only: t_Interpolator
    use m_scatter, only: t_Scatter
    use m_particle_boundary, only: t_ParticleBoundary
    use m_field_solver, only: t_FieldSolver
    use m_mpi_fft_solver, only: t_MPIFFTSolver3d
    implicit none
    private
    public :: t_OhpicSimulation, ohpic_check
    public :: t_ParticleMoverHip, create_particle_mover_hip
    public :: t_InterpolatorHip, t_ScatterHip
    public :: t_ParticleBoundariesHip, new_ParticleBoundariesHip
    public :: t_PoissonFieldSolverHip, t_FFTSolverHip, new_FFTSolverHip
    public :: t_Hdf5ForOhfieldHip, new_Hdf5ForOhfieldHip

    !> One device-resident simulation (an ohpic_ctx).
    type :: t_OhpicSimulation
        type(c_ptr) :: ctx = c_null_ptr
        integer :: n(3) = 0
        integer :: nspecies = 0
        integer :: device = 0
    contains
        procedure :: init => sim_init
        procedure :: destroy => sim_destroy
        procedure :: load => sim_load
        procedure :: store => sim_store
        procedure :: count => sim_count
        procedure :: seed_uniform_maxwell => sim_seed_uniform_maxwell
        procedure :: step => sim_step
        procedure :: get_field => sim_get_field
        procedure :: put_field => sim_put_field
        procedure :: diagnostics => sim_diagnostics
        procedure :: energy => sim_energy
    end type

    type, extends(t_ParticleMover) :: t_ParticleMoverHip
        integer :: device = 0
    contains
        procedure :: move => hip_move
        procedure :: move_all => hip_move_all
    end type

    type, extends(t_Interpolator) :: t_InterpolatorHip
        integer :: device = 0
        integer :: weight_order = OHPIC_WEIGHTS_REFERENCE
    contains
        procedure :: interp => hip_interp
    end type

    type, extends(t_Scatter) :: t_ScatterHip
        integer :: device = 0
        integer :: weight_order = OHPIC_WEIGHTS_REFERENCE
    contains
        procedure :: scatter => hip_scatter
        procedure :: scatter_all => hip_scatter_all
    end type

    type, extends(t_ParticleBoundary) :: t_ParticleBoundariesHip
        integer :: n(3) = 0
        integer :: device = 0
    contains
        procedure :: apply => hip_apply
        procedure :: apply_all => hip_apply_all
    end type

    type, extends(t_FieldSolver) :: t_PoissonFieldSolverHip
        type(t_OhpicSimulation), pointer :: sim => null()
    contains
        procedure :: solve => hip_solve
    end type

    type, extends(t_MPIFFTSolver3d) :: t_FFTSolverHip
        integer :: device = 0
    contains
        procedure :: forward => hip_fft_forward
        procedure :: backward => hip_fft_backward
    end type

    !> t_Hdf5ForOhfield (hdf5_for_ohfield.f90:11-22) over ohpic_snapshot_*: one
    !> HDF5 file and group, one dataset of global_shape per write.
    type :: t_Hdf5ForOhfieldHip
        type(c_ptr) :: snap = c_null_ptr
        character(:), allocatable :: filename
        character(:), allocatable :: group_name
        integer(kind=8) :: global_shape(3) = 0
        integer(kind=8) :: global_offset(3) = 0
    contains
        procedure :: write => h5_write
        procedure :: write_device => h5_write_device
        procedure :: close => h5_close
    end type

contains

    subroutine ohpic_check(rc, ctx)
        integer(c_int), intent(in) :: rc
        type(c_ptr), intent(in), optional :: ctx
        type(c_ptr) :: msg
        character(kind=c_char), pointer :: chars(:)
        character(len=512) :: text
        integer :: i
        if (rc == 0) return
        if (present(ctx)) then
            msg = ohpic_last_error(ctx)
        else
            msg = ohpic_last_error(c_null_ptr)
        end if
        text = ''
        if (c_associated(msg)) then
            call c_f_pointer(msg, chars, [512])
            do i = 1, 512
                if (chars(i) == c_null_char) exit
                text(i:i) = chars(i)
            end do
        end if
        write (*, '(a,i0,a,a)') 'ohpic error ', rc, ': ', trim(text)
        error stop 1
    end subroutine

    ! ----------------------------------------------------------------- simulation
    subroutine sim_init(self, n, qm, q_macro, capacity, dt, b0, precision, weight_order, &
                        track_ids, device)
        class(t_OhpicSimulation), intent(inout) :: self
        integer, intent(in) :: n(3)
        double precision, intent(in) :: qm(:), q_macro(:)
        integer(c_int64_t), intent(in) :: capacity(:)
        double precision, intent(in) :: dt, b0(3)
        integer, intent(in), optional :: precision, weight_order, device
        logical, intent(in), optional :: track_ids
        type(ohpic_config) :: cfg
        integer :: s
        call ohpic_check(ohpic_config_default(cfg))
        cfg%nx = n(1)
        cfg%ny = n(2)
        cfg%nz = n(3)
        cfg%nspecies = size(qm)
        do s = 1, size(qm)
            cfg%qm(s) = qm(s)
            cfg%q_macro(s) = q_macro(s)
            cfg%capacity(s) = capacity(s)
        end do
        cfg%dt = dt
        cfg%b0 = b0
        if (present(precision)) cfg%precision = precision
        if (present(weight_order)) cfg%weight_order = weight_order
        if (present(device)) cfg%device = device
        if (present(track_ids)) then
            if (track_ids) cfg%track_ids = 1
        end if
        call ohpic_check(ohpic_create(cfg, self%ctx))
        self%n = n
        self%nspecies = size(qm)
        self%device = cfg%device
    end subroutine

    subroutine sim_destroy(self)
        class(t_OhpicSimulation), intent(inout) :: self
        if (c_associated(self%ctx)) call ohpic_check(ohpic_destroy(self%ctx))
        self%ctx = c_null_ptr
    end subroutine

    !> Replace the particles with oh_particle records (spec selects the species).
    subroutine sim_load(self, particles)
        class(t_OhpicSimulation), intent(inout) :: self
        type(oh_particle), intent(in) :: particles(:)
        call ohpic_check(ohpic_load_oh_particles(self%ctx, int(size(particles), c_int64_t), &
                                                 particles), self%ctx)
    end subroutine

    integer(c_int64_t) function sim_count(self)
        class(t_OhpicSimulation), intent(inout) :: self
        integer(c_int64_t) :: m
        integer :: s
        sim_count = 0
        do s = 0, self%nspecies - 1
            call ohpic_check(ohpic_particle_count(self%ctx, s, m), self%ctx)
            sim_count = sim_count + m
        end do
    end function

    !> Copy every particle out as oh_particle records; n returns the count.
    subroutine sim_store(self, particles, n)
        class(t_OhpicSimulation), intent(inout) :: self
        type(oh_particle), intent(inout) :: particles(:)
        integer(c_int64_t), intent(out) :: n
        call ohpic_check(ohpic_store_oh_particles(self%ctx, int(size(particles), c_int64_t), n, &
                                                  particles), self%ctx)
    end subroutine

    subroutine sim_seed_uniform_maxwell(self, np, vth, seeds)
        class(t_OhpicSimulation), intent(inout) :: self
        integer(c_int64_t), intent(in) :: np(:)
        double precision, intent(in) :: vth(:)
        integer(c_int64_t), intent(in) :: seeds(2)
        call ohpic_check(ohpic_seed_uniform_maxwell(self%ctx, seeds(1), seeds(2), np, vth), &
                         self%ctx)
    end subroutine

    subroutine sim_step(self, nsteps)
        class(t_OhpicSimulation), intent(inout) :: self
        integer, intent(in) :: nsteps
        call ohpic_check(ohpic_step(self%ctx, int(nsteps, c_int)), self%ctx)
    end subroutine

    !> Device field -> field%values(comp, ...) including halos (periodic fill).
    subroutine sim_get_field(self, which, field, comp)
        class(t_OhpicSimulation), intent(inout) :: self
        integer, intent(in) :: which
        class(t_OhField), intent(inout) :: field
        integer, intent(in), optional :: comp
        integer(c_int32_t) :: lb(3), ub(3)
        integer :: a, c
        c = 1
        if (present(comp)) c = comp
        do a = 1, 3
            lb(a) = lbound(field%values, a + 1)
            ub(a) = ubound(field%values, a + 1)
        end do
        call ohpic_check(ohpic_get_field(self%ctx, int(which, c_int), field%values(:, :, :, :, 1), &
                                         int(field%nelements, c_int), lb, ub, int(c - 1, c_int)), &
                         self%ctx)
    end subroutine

    !> field%values(comp, 0:n-1, ...) -> device field.
    subroutine sim_put_field(self, which, field, comp)
        class(t_OhpicSimulation), intent(inout) :: self
        integer, intent(in) :: which
        class(t_OhField), intent(in) :: field
        integer, intent(in), optional :: comp
        integer(c_int32_t) :: lb(3), ub(3)
        integer :: a, c
        c = 1
        if (present(comp)) c = comp
        do a = 1, 3
            lb(a) = lbound(field%values, a + 1)
            ub(a) = ubound(field%values, a + 1)
        end do
        call ohpic_check(ohpic_put_field(self%ctx, int(which, c_int), field%values(:, :, :, :, 1), &
                                         int(field%nelements, c_int), lb, ub, int(c - 1, c_int)), &
                         self%ctx)
    end subroutine

    subroutine sim_diagnostics(self, diag)
        class(t_OhpicSimulation), intent(inout) :: self
        type(ohpic_diag), intent(out) :: diag
        call ohpic_check(ohpic_diagnostics(self%ctx, diag), self%ctx)
    end subroutine

    ! [kinetic, field, total, net charge] of the last step
    function sim_energy(self) result(e)
        class(t_OhpicSimulation), intent(inout) :: self
        real(c_double) :: e(4)
        call ohpic_check(ohpic_energy(self%ctx, e), self%ctx)
    end function

    ! ---------------------------------------------------------------- mover
    function create_particle_mover_hip(name, device) result(mover)
        character(*), intent(in) :: name
        integer, intent(in), optional :: device
        class(t_ParticleMover), allocatable :: mover
        type(t_ParticleMoverHip) :: boris
        select case (name)
        case ('boris')
            if (present(device)) boris%device = device
            allocate (mover, source=boris)
        case default
            write (*, '(a)') 'ohpic error: unknown particle mover '//name
            error stop 1
        end select
    end function

    !> t_ParticleMoverBoris%move (particle_mover_boris.f90:21-64) for one particle.
    subroutine hip_move(self, particle, qm, eb, dt)
        class(t_ParticleMoverHip), intent(in) :: self
        type(oh_particle), intent(inout) :: particle
        double precision, intent(in) :: qm
        double precision, intent(in) :: eb(6)
        double precision :: dt
        real(c_double) :: x(1), y(1), z(1), vx(1), vy(1), vz(1)
        x(1) = particle%x
        y(1) = particle%y
        z(1) = particle%z
        vx(1) = particle%vx
        vy(1) = particle%vy
        vz(1) = particle%vz
        call ohpic_check(ohpic_boris_batch(int(self%device, c_int), 1_c_int64_t, x, y, z, vx, &
                                           vy, vz, eb, qm, dt))
        particle%x = x(1)
        particle%y = y(1)
        particle%z = z(1)
        particle%vx = vx(1)
        particle%vy = vy(1)
        particle%vz = vz(1)
    end subroutine

    !> Gather + Boris for every particle of the simulation (no boundary).
    subroutine hip_move_all(self, sim)
        class(t_ParticleMoverHip), intent(in) :: self
        class(t_OhpicSimulation), intent(inout) :: sim
        call ohpic_check(ohpic_gather_push(sim%ctx), sim%ctx)
    end subroutine

    ! ----------------------------------------------------------- interpolator
    !> t_LinearInterpolator%interp (linear_interpolator.f90:22-69).
    function hip_interp(self, particle, field, ps) result(ret)
        class(t_InterpolatorHip), intent(in) :: self
        type(oh_particle), intent(in) :: particle
        class(t_OhField), intent(in) :: field
        integer, intent(in) :: ps
        double precision :: ret(field%nelements)
        real(c_double) :: pos(3)
        integer(c_int32_t) :: lb(3), ub(3)
        integer :: a
        pos = field%to_local_position([particle%x, particle%y, particle%z], ps)
        do a = 1, 3
            lb(a) = lbound(field%values, a + 1)
            ub(a) = ubound(field%values, a + 1)
        end do
        call ohpic_check(ohpic_interp_batch(int(self%device, c_int), 1_c_int64_t, pos, &
                                            field%values(:, :, :, :, ps), &
                                            int(field%nelements, c_int), lb, ub, &
                                            int(self%weight_order, c_int), ret))
    end function

    ! ---------------------------------------------------------------- scatter
    !> t_LinearScatter%scatter (linear_scatter.f90:22-55) for one particle.
    subroutine hip_scatter(self, particle, ohfield, amount, ps)
        class(t_ScatterHip), intent(in) :: self
        type(oh_particle), intent(in) :: particle
        class(t_OhField), intent(inout) :: ohfield
        double precision, intent(in) :: amount(ohfield%nelements)
        integer, intent(in) :: ps
        real(c_double) :: pos(3)
        integer(c_int32_t) :: lb(3), ub(3)
        integer :: a
        pos = ohfield%to_local_position([particle%x, particle%y, particle%z], ps)
        do a = 1, 3
            lb(a) = lbound(ohfield%values, a + 1)
            ub(a) = ubound(ohfield%values, a + 1)
        end do
        call ohpic_check(ohpic_scatter_batch(int(self%device, c_int), 1_c_int64_t, pos, &
                                             ohfield%values(:, :, :, :, ps), &
                                             int(ohfield%nelements, c_int), lb, ub, amount, &
                                             int(self%weight_order, c_int)))
    end subroutine

    !> rho := sum over all particles (mainstep deposit, ohhelppic3d.f90:311-320);
    !> the charge also lands in rho%values when rho is given.
    subroutine hip_scatter_all(self, sim, rho)
        class(t_ScatterHip), intent(in) :: self
        class(t_OhpicSimulation), intent(inout) :: sim
        class(t_OhField), intent(inout), optional :: rho
        call ohpic_check(ohpic_deposit(sim%ctx), sim%ctx)
        if (present(rho)) call sim%get_field(OHPIC_FIELD_RHO, rho)
    end subroutine

    ! ------------------------------------------------------------- boundaries
    function new_ParticleBoundariesHip(nx, ny, nz, device) result(obj)
        integer, intent(in) :: nx, ny, nz
        integer, intent(in), optional :: device
        type(t_ParticleBoundariesHip) :: obj
        obj%n = [nx, ny, nz]
        if (present(device)) obj%device = device
    end function

    !> t_ParticleBoundaries%apply (particle_boundaries.f90:34-43): periodic
    !> wrap of x, y and z (periodic_particle_boundary.f90:42-53).
    subroutine hip_apply(self, particle, dt)
        class(t_ParticleBoundariesHip), intent(in) :: self
        type(oh_particle), intent(inout) :: particle
        double precision, intent(in) :: dt
        real(c_double) :: p(1)
        p(1) = particle%x
        call ohpic_check(ohpic_wrap_batch(int(self%device, c_int), 1_c_int64_t, p, &
                                          int(self%n(1), c_int32_t), 1_c_int))
        particle%x = p(1)
        p(1) = particle%y
        call ohpic_check(ohpic_wrap_batch(int(self%device, c_int), 1_c_int64_t, p, &
                                          int(self%n(2), c_int32_t), 1_c_int))
        particle%y = p(1)
        p(1) = particle%z
        call ohpic_check(ohpic_wrap_batch(int(self%device, c_int), 1_c_int64_t, p, &
                                          int(self%n(3), c_int32_t), 1_c_int))
        particle%z = p(1)
    end subroutine

    !> Wrap (applied twice, as mainstep does) and re-bin every particle.
    subroutine hip_apply_all(self, sim)
        class(t_ParticleBoundariesHip), intent(in) :: self
        class(t_OhpicSimulation), intent(inout) :: sim
        if (any(self%n /= sim%n)) then
            write (*, '(a)') 'ohpic error: boundary extents differ from the simulation grid'
            error stop 1
        end if
        call ohpic_check(ohpic_boundary(sim%ctx), sim%ctx)
    end subroutine

    ! ------------------------------------------------------------ field solve
    !> t_PoissonFieldSolver%solve (poisson_field_solver.f90:90-175), periodic:
    !> rho%values -> device -> phi, E -> phi%values and eb%values (E in
    !> components 1:3, the uniform B0 in 4:6), halos filled periodically.
    subroutine hip_solve(self, rho, aj, eb, phi, ohhelp)
        class(t_PoissonFieldSolverHip), intent(in) :: self
        class(t_OhField), intent(in) :: rho
        class(t_OhField), intent(in) :: aj
        class(t_OhField), intent(inout) :: eb
        class(t_OhField), intent(inout) :: phi
        class(t_OhHelp), intent(inout) :: ohhelp
        if (.not. associated(self%sim)) then
            write (*, '(a)') 'ohpic error: t_PoissonFieldSolverHip has no simulation'
            error stop 1
        end if
        call self%sim%put_field(OHPIC_FIELD_RHO, rho)
        call ohpic_check(ohpic_field_solve(self%sim%ctx), self%sim%ctx)
        call self%sim%get_field(OHPIC_FIELD_PHI, phi)
        call self%sim%get_field(OHPIC_FIELD_EB, eb)
    end subroutine

    ! -------------------------------------------------------------------- FFT
    !> Periodic FFT over an n(1) x n(2) x n(3) block (logical block 0:n-1).
    function new_FFTSolverHip(n, device) result(obj)
        integer, intent(in) :: n(3)
        integer, intent(in), optional :: device
        type(t_FFTSolverHip) :: obj
        obj%boundary_types = 0
        obj%local_block = new_Block([0, 0, 0], n - 1)
        obj%logical_block = obj%local_block
        obj%global_block = new_Block([0, 0, 0], n)
        if (present(device)) obj%device = device
    end function

    !> R2HC along every axis (FFTW halfcomplex layout).
    subroutine hip_fft_forward(self, in, out)
        class(t_FFTSolverHip), intent(in) :: self
        double precision, intent(in) :: in(self%local_block%start(1):self%local_block%end(1), &
                                           self%local_block%start(2):self%local_block%end(2), &
                                           self%local_block%start(3):self%local_block%end(3))
        double precision, intent(out) :: out(self%local_block%start(1):self%local_block%end(1), &
                                             self%local_block%start(2):self%local_block%end(2), &
                                             self%local_block%start(3):self%local_block%end(3))
        call ohpic_check(ohpic_fft3d_r2r(int(self%device, c_int), &
                                         int(self%local_block%sizes(1), c_int), &
                                         int(self%local_block%sizes(2), c_int), &
                                         int(self%local_block%sizes(3), c_int), 0_c_int, in, out))
    end subroutine

    !> HC2R along every axis divided by the number of nodes
    !> (mpi_fftw3_solver.f90:223-236).
    subroutine hip_fft_backward(self, in, out)
        class(t_FFTSolverHip), intent(in) :: self
        double precision, intent(in) :: in(self%local_block%start(1):self%local_block%end(1), &
                                           self%local_block%start(2):self%local_block%end(2), &
                                           self%local_block%start(3):self%local_block%end(3))
        double precision, intent(out) :: out(self%local_block%start(1):self%local_block%end(1), &
                                             self%local_block%start(2):self%local_block%end(2), &
                                             self%local_block%start(3):self%local_block%end(3))
        call ohpic_check(ohpic_fft3d_r2r(int(self%device, c_int), &
                                         int(self%local_block%sizes(1), c_int), &
                                         int(self%local_block%sizes(2), c_int), &
                                         int(self%local_block%sizes(3), c_int), 1_c_int, in, out))
    end subroutine

    ! ----------------------------------------------------------------- snapshots
    !> new_Hdf5ForOhfield (hdf5_for_ohfield.f90:27-43); comm is accepted for
    !> the same call shape and unused (one rank writes the whole box).
    function new_Hdf5ForOhfieldHip(filename, group_name, global_shape, global_offset, comm) &
        result(obj)
        character(*), intent(in) :: filename
        character(*), intent(in) :: group_name
        integer(kind=8), intent(in) :: global_shape(3)
        integer(kind=8), intent(in) :: global_offset(3)
        integer, intent(in), optional :: comm
        type(t_Hdf5ForOhfieldHip) :: obj
        obj%filename = filename
        obj%group_name = group_name
        obj%global_shape = global_shape
        obj%global_offset = global_offset
        call ohpic_check(ohpic_snapshot_open(filename//c_null_char, group_name//c_null_char, &
                                             int(global_shape, c_int64_t), &
                                             int(global_offset, c_int64_t), obj%snap))
    end function

    !> hdf5ForOhfield_write (hdf5_for_ohfield.f90:45-76): the host block
    !> values(1, 0:xu-xl, 0:yu-yl, 0:zu-zl, ps) at offset (xl, yl, zl).
    subroutine h5_write(self, dataset_name, ohfield, ps)
        class(t_Hdf5ForOhfieldHip), intent(in) :: self
        character(*), intent(in) :: dataset_name
        class(t_OhField), intent(in) :: ohfield
        integer, intent(in) :: ps
        integer :: xl, xu, yl, yu, zl, zu
        integer(c_int64_t) :: shp(3), off(3)
        double precision, allocatable :: block(:, :, :)
        xl = ohfield%subdomain_range(1, 1, ps)
        xu = ohfield%subdomain_range(2, 1, ps)
        yl = ohfield%subdomain_range(1, 2, ps)
        yu = ohfield%subdomain_range(2, 2, ps)
        zl = ohfield%subdomain_range(1, 3, ps)
        zu = ohfield%subdomain_range(2, 3, ps)
        shp = [xu - xl + 1, yu - yl + 1, zu - zl + 1]
        off = [xl, yl, zl]
        block = ohfield%values(1, 0:xu - xl, 0:yu - yl, 0:zu - zl, ps)
        call ohpic_check(ohpic_snapshot_write_host(self%snap, dataset_name//c_null_char, block, &
                                                   shp, off))
    end subroutine

    !> The same dataset written from the device field `which` of a simulation
    !> (nodes 0..n per axis, node n = node 0), without a host t_OhField.
    subroutine h5_write_device(self, dataset_name, sim, which)
        class(t_Hdf5ForOhfieldHip), intent(in) :: self
        character(*), intent(in) :: dataset_name
        class(t_OhpicSimulation), intent(in) :: sim
        integer, intent(in) :: which
        call ohpic_check(ohpic_snapshot_write(self%snap, dataset_name//c_null_char, sim%ctx, &
                                              int(which, c_int)))
    end subroutine

    !> hdf5ForOhfield_close (hdf5_for_ohfield.f90:78-86).
    subroutine h5_close(self)
        class(t_Hdf5ForOhfieldHip), intent(inout) :: self
        call ohpic_check(ohpic_snapshot_close(self%snap))
        self%snap = c_null_ptr
    end subroutine

end module
