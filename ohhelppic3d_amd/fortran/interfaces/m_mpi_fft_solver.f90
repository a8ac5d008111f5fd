!> Abstract 3-D real FFT over a block (interface of
!> src/field_solver/fft/mpi_fft_solver.f90:8-40).
module m_mpi_fft_solver
    use m_block, only: t_Block
    implicit none
    private
    public t_MPIFFTSolver3d

    type, abstract :: t_MPIFFTSolver3d
        integer :: boundary_types(3) = 0
        type(t_Block) :: local_block
        type(t_Block) :: logical_block
        type(t_Block) :: global_block
    contains
        procedure(mpiFFTSolver3d_transform), deferred :: forward
        procedure(mpiFFTSolver3d_transform), deferred :: backward
    end type

    abstract interface
        subroutine mpiFFTSolver3d_transform(self, in, out)
            import t_MPIFFTSolver3d
            class(t_MPIFFTSolver3d), intent(in) :: self
            double precision, intent(in) :: in(self%local_block%start(1):self%local_block%end(1), &
                                               self%local_block%start(2):self%local_block%end(2), &
                                               self%local_block%start(3):self%local_block%end(3))
            double precision, intent(out) :: out(self%local_block%start(1):self%local_block%end(1), &
                                                 self%local_block%start(2):self%local_block%end(2), &
                                                 self%local_block%start(3):self%local_block%end(3))
        end subroutine
    end interface
end module
