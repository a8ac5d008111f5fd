!> Abstract particle boundary (interface of src/particle_boundary/particle_boundary.f90:8-21).
module m_particle_boundary
    use oh_type, only: oh_particle
    implicit none
    private
    public t_ParticleBoundary

    type, abstract :: t_ParticleBoundary
    contains
        procedure(particleBoundary_apply), deferred :: apply
    end type

    abstract interface
        subroutine particleBoundary_apply(self, particle, dt)
            import t_ParticleBoundary, oh_particle
            class(t_ParticleBoundary), intent(in) :: self
            type(oh_particle), intent(inout) :: particle
            double precision, intent(in) :: dt
        end subroutine
    end interface
end module
