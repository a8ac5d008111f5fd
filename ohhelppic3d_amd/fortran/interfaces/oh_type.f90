!> Single-GPU stand-in for OhHelp's oh_type module (not vendored with the
!> reference).  oh_particle carries the fields the reference uses
!> (particle_injector_with_distribution.f90:59-69, ohhelp.f90:423-426) in a
!> C-interoperable layout identical to ohpic_oh_particle in include/ohpic.h.
module oh_type
    use iso_c_binding, only: c_double, c_int64_t, c_int32_t
    implicit none
    private
    public oh_particle

    type, bind(c) :: oh_particle
        real(c_double) :: x, y, z
        real(c_double) :: vx, vy, vz
        integer(c_int64_t) :: pid
        integer(c_int32_t) :: nid
        integer(c_int32_t) :: spec
    end type
end module
