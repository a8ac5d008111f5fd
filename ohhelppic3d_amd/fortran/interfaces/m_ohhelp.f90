!> Single-GPU stand-in for the reference's OhHelp wrapper t_OhHelp
!> (src/ohhelp/ohhelp.f90:35-86): one rank owns the whole periodic domain, it
!> is always in primary mode, broadcast/reduce are identities and
!> exchange_borders is a periodic ghost fill.
module m_ohhelp
    use m_ohfield, only: t_OhField
    implicit none
    private
    public t_OhHelp

    type :: t_OhHelp
        integer :: n(3) = 0
    contains
        procedure :: is_primary_mode => ohhelp_is_primary_mode
        procedure :: is_secondary_mode => ohhelp_is_secondary_mode
        procedure :: exchange_borders => ohhelp_exchange_borders
        procedure :: broadcast_field => ohhelp_noop
        procedure :: reduce_field => ohhelp_noop
    end type
contains
    logical function ohhelp_is_primary_mode(self)
        class(t_OhHelp), intent(in) :: self
        ohhelp_is_primary_mode = .true.
    end function

    logical function ohhelp_is_secondary_mode(self)
        class(t_OhHelp), intent(in) :: self
        ohhelp_is_secondary_mode = .false.
    end function

    subroutine ohhelp_noop(self, field)
        class(t_OhHelp), intent(inout) :: self
        class(t_OhField), intent(inout) :: field
    end subroutine

    !> Copy interior nodes [0, n) into every halo node (periodic images).
    subroutine ohhelp_exchange_borders(self, field)
        class(t_OhHelp), intent(inout) :: self
        class(t_OhField), intent(inout) :: field
        integer :: i, j, k, lo(3), hi(3)
        do i = 1, 3
            lo(i) = lbound(field%values, i + 1)
            hi(i) = ubound(field%values, i + 1)
        end do
        do k = lo(3), hi(3)
            do j = lo(2), hi(2)
                do i = lo(1), hi(1)
                    field%values(:, i, j, k, 1) = field%values(:, modulo(i, self%n(1)), &
                                                               modulo(j, self%n(2)), &
                                                               modulo(k, self%n(3)), 1)
                end do
            end do
        end do
    end subroutine
end module
