!> Abstract particle mover (interface of src/particle_mover/particle_mover.f90:8-25).
module m_particle_mover
    use oh_type, only: oh_particle
    implicit none
    private
    public t_ParticleMover

    type, abstract :: t_ParticleMover
    contains
        procedure(particleMover_move), deferred :: move
    end type

    abstract interface
        subroutine particleMover_move(self, particle, qm, eb, dt)
            import t_ParticleMover, oh_particle
            class(t_ParticleMover), intent(in) :: self
            type(oh_particle), intent(inout) :: particle
            double precision, intent(in) :: qm
            double precision, intent(in) :: eb(6)
            double precision :: dt
        end subroutine
    end interface
end module
