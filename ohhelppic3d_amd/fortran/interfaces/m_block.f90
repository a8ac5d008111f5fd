!> Index box with inclusive bounds (the reference's t_Block,
!> src/field_solver/fft/block/block.f90:11-32): sizes = max(end - start + 1, 0).
module m_block
    implicit none
    private
    public t_Block, new_Block

    type :: t_Block
        integer :: start(3) = 0
        integer :: end(3) = -1
        integer :: sizes(3) = 0
    end type
contains
    function new_Block(start, end) result(obj)
        integer, intent(in) :: start(3), end(3)
        type(t_Block) :: obj
        obj%start = start
        obj%end = end
        obj%sizes = max(end - start + 1, 0)
    end function
end module
