!> Abstract charge scatter (interface of src/scatter/scatter.f90:9-25).
module m_scatter
    use oh_type, only: oh_particle
    use m_ohfield, only: t_OhField
    implicit none
    private
    public t_Scatter

    type, abstract :: t_Scatter
    contains
        procedure(scatter_scatter), deferred :: scatter
    end type

    abstract interface
        subroutine scatter_scatter(self, particle, ohfield, amount, ps)
            import t_Scatter, oh_particle, t_OhField
            class(t_Scatter), intent(in) :: self
            type(oh_particle), intent(in) :: particle
            class(t_OhField), intent(inout) :: ohfield
            double precision, intent(in) :: amount(ohfield%nelements)
            integer, intent(in) :: ps
        end subroutine
    end interface
end module
