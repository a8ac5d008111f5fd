!> Single-GPU stand-in for the reference's field container t_OhField
!> (src/ohhelp/ohfield.f90:53-72): values(nelements, lx:ux, ly:uy, lz:uz,
!> nfields) over one periodic subdomain whose origin is 0, so
!> to_local_position (ohfield.f90:199-206) is the identity.
module m_ohfield
    implicit none
    private
    public t_OhField, new_OhField

    type :: t_OhField
        double precision, allocatable :: values(:, :, :, :, :)
        integer :: nelements = 0
        integer :: nfields = 1
        integer :: subdomain_range(2, 3, 2) = 0
    contains
        procedure :: to_local_position => ohfield_to_local_position
        procedure :: to_local_index => ohfield_to_local_index
    end type
contains
    !> Field of nelements components on an n(1) x n(2) x n(3) periodic grid
    !> with halo widths lext below and uext above every axis.
    function new_OhField(nelements, n, lext, uext) result(obj)
        integer, intent(in) :: nelements, n(3), lext, uext
        type(t_OhField) :: obj
        integer :: a
        obj%nelements = nelements
        obj%nfields = 1
        do a = 1, 3
            obj%subdomain_range(1, a, :) = 0
            obj%subdomain_range(2, a, :) = n(a)
        end do
        allocate (obj%values(nelements, -lext:n(1) - 1 + uext, -lext:n(2) - 1 + uext, &
                             -lext:n(3) - 1 + uext, 1))
        obj%values = 0d0
    end function

    function ohfield_to_local_position(self, global_position, ps) result(ret)
        class(t_OhField), intent(in) :: self
        double precision, intent(in) :: global_position(3)
        integer, intent(in) :: ps
        double precision :: ret(3)
        ret = global_position - self%subdomain_range(1, :, ps)
    end function

    function ohfield_to_local_index(self, global_index, ps) result(ret)
        class(t_OhField), intent(in) :: self
        integer, intent(in) :: global_index(3)
        integer, intent(in) :: ps
        integer :: ret(3)
        ret = global_index - self%subdomain_range(1, :, ps)
    end function
end module
