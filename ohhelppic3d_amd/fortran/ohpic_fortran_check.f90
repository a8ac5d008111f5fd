!> Drives the GPU implementation through the reference's Fortran interfaces.
!>
!> usage: ohpic_fortran_check <input.bin> <output.bin> [<snapshot.h5>]
!> input (stream): int32 n(3), nspecies, nsteps, kcheck; int64 np(nspecies);
!>                 float64 qm(nspecies), q(nspecies), dt, b0(3); oh_particle records
!> output (stream): see the write statements below, in order.
!>
!> Part 1 runs `nsteps` of mainstep (ohhelppic3d.f90:311-358) through
!> polymorphic t_Scatter / t_FieldSolver / t_ParticleMover / t_ParticleBoundary
!> objects on host t_OhField arrays; part 2 calls the per-particle bindings
!> (interp, move, apply twice, scatter) on the first kcheck input particles;
!> part 3 transforms the reference's FFT impulse (1 x 1 x 63); part 4 runs the
!> fused step on a second simulation from the same input; part 5 (with a third
!> argument) writes phi of part 1 (host t_OhField) and of part 4 (device) as
!> datasets '1' and '2' of group 'phisp' (ohhelppic3d.f90:79-99).
program ohpic_fortran_check
    use iso_c_binding
    use oh_type, only: oh_particle
    use m_ohfield, only: t_OhField, new_OhField
    use m_ohhelp, only: t_OhHelp
    use m_particle_mover, only: t_ParticleMover
    use m_interpolator, only: t_Interpolator
    use m_scatter, only: t_Scatter
    use m_particle_boundary, only: t_ParticleBoundary
    use m_field_solver, only: t_FieldSolver
    use m_mpi_fft_solver, only: t_MPIFFTSolver3d
    use m_ohpic
    use m_ohpic_c, only: OHPIC_FIELD_PHI
    implicit none

    character(len=1024) :: inpath, outpath, h5path
    type(t_Hdf5ForOhfieldHip) :: h5
    integer(c_int32_t) :: hdr(6)
    integer :: n(3), nsp, nsteps, kcheck, u, i, s, istep
    integer(c_int64_t), allocatable :: np(:)
    double precision, allocatable :: qm(:), q(:)
    double precision :: dt, b0(3)
    type(oh_particle), allocatable :: parts(:), outp(:), probe(:)
    integer(c_int64_t) :: ntot, nout
    type(t_OhpicSimulation), target :: sim, sim2
    type(t_OhField) :: rho, phi, eb, aj, rho2
    type(t_OhHelp) :: ohhelp
    class(t_ParticleMover), allocatable :: mover
    class(t_Interpolator), allocatable :: interp
    class(t_Scatter), allocatable :: scatter
    class(t_ParticleBoundary), allocatable :: boundaries
    class(t_FieldSolver), allocatable :: solver
    class(t_MPIFFTSolver3d), allocatable :: fft
    double precision, allocatable :: ebi(:, :)
    double precision :: impulse(1, 1, 63), spectrum(1, 1, 63)
    double precision :: en(4)

    call get_command_argument(1, inpath)
    call get_command_argument(2, outpath)
    h5path = ''
    if (command_argument_count() >= 3) call get_command_argument(3, h5path)
    open (newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old')
    read (u) hdr
    n = hdr(1:3)
    nsp = hdr(4)
    nsteps = hdr(5)
    kcheck = hdr(6)
    allocate (np(nsp), qm(nsp), q(nsp))
    read (u) np, qm, q, dt, b0
    ntot = sum(np)
    allocate (parts(ntot), outp(ntot))
    read (u) parts
    close (u)

    ! ---- part 1: mainstep through the polymorphic objects ------------------
    call sim%init(n, qm, q, np, dt, b0, track_ids=.true.)
    call sim%load(parts)
    rho = new_OhField(1, n, 0, 1)
    aj = new_OhField(3, n, 0, 0)
    phi = new_OhField(1, n, 1, 2)
    eb = new_OhField(6, n, 1, 2)
    ohhelp%n = n
    allocate (t_ScatterHip :: scatter)
    allocate (t_PoissonFieldSolverHip :: solver)
    select type (solver)
    type is (t_PoissonFieldSolverHip)
        solver%sim => sim
    end select
    mover = create_particle_mover_hip('boris')
    allocate (boundaries, source=new_ParticleBoundariesHip(n(1), n(2), n(3)))
    allocate (t_InterpolatorHip :: interp)
    do istep = 1, nsteps
        select type (scatter)
        type is (t_ScatterHip)
            call scatter%scatter_all(sim, rho)
        end select
        call solver%solve(rho, aj, eb, phi, ohhelp)
        select type (mover)
        type is (t_ParticleMoverHip)
            call mover%move_all(sim)
        end select
        select type (boundaries)
        type is (t_ParticleBoundariesHip)
            call boundaries%apply_all(sim)
        end select
    end do
    call sim%store(outp, nout)

    ! ---- part 2: per-particle bindings on the input's first kcheck particles
    allocate (probe(kcheck), ebi(6, kcheck))
    probe = parts(1:kcheck)
    do i = 1, kcheck
        ebi(:, i) = interp%interp(probe(i), eb, 1)
        call mover%move(probe(i), qm(probe(i)%spec), ebi(:, i), dt)
        call boundaries%apply(probe(i), dt)
        call boundaries%apply(probe(i), dt)
    end do
    rho2 = new_OhField(1, n, 0, 1)
    do i = 1, kcheck
        s = parts(i)%spec
        call scatter%scatter(parts(i), rho2, [q(s)], 1)
    end do

    ! ---- part 3: FFT impulse response (test_mpi_fftw3_solver.f90:53-97) -----
    allocate (fft, source=new_FFTSolverHip([1, 1, 63]))
    impulse = 0d0
    impulse(1, 1, 2) = 1d0
    call fft%forward(impulse, spectrum)

    ! ---- part 4: the fused step from the same input -------------------------
    call sim2%init(n, qm, q, np, dt, b0, track_ids=.true.)
    call sim2%load(parts)
    call sim2%step(nsteps)
    ! energy history entry of the last step: kinetic + field = total
    en = sim2%energy()
    if (.not. (en(1) > 0d0 .and. abs(en(1) + en(2) - en(3)) <= 1d-12 * abs(en(3)))) &
        error stop 'ohpic_energy: inconsistent kinetic, field and total energy'

    ! ---- part 5: snapshots in the reference's phisp layout --------------------
    if (len_trim(h5path) > 0) then
        h5 = new_Hdf5ForOhfieldHip(trim(h5path), 'phisp', int(n + 1, kind=8), &
                                   int([0, 0, 0], kind=8))
        call h5%write('1', phi, 1)
        call h5%write_device('2', sim2, OHPIC_FIELD_PHI)
        call h5%close()
    end if

    open (newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
    write (u) nout
    write (u) outp(1:nout)
    write (u) rho%values
    write (u) phi%values
    write (u) eb%values
    write (u) probe
    write (u) ebi
    write (u) rho2%values
    write (u) spectrum
    call sim2%store(outp, nout)
    write (u) nout
    write (u) outp(1:nout)
    close (u)
    call sim%destroy()
    call sim2%destroy()
end program
