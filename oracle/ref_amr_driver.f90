! ref_amr_driver.f90 -- C-callable entry points into the reference's regrid
! averaging and time interpolation (fc2d_geoclaw_average_fort.f,
! fc2d_geoclaw_timeinterp_fort.f), compiled by oracle/Makefile into
! oracle/_ref/libgfref_amr.so.
!
! TEST INFRASTRUCTURE ONLY.  The same source files also hold the face and
! corner ghost routines, which call ForestClaw's index transforms (not in the
! reference tree); those references stay unresolved, so the library is loaded
! with lazy binding and only the routines below are ever called.

module gfref_amr_driver
    use iso_c_binding
    implicit none
contains

    subroutine gfref_amr_set_drytol(drytol) bind(C, name="gfref_amr_set_drytol")
        use geoclaw_module, only: dry_tolerance
        real(c_double), value :: drytol
        dry_tolerance = drytol
    end subroutine gfref_amr_set_drytol

    subroutine gfref_average2coarse(mx, my, mbc, maux, qcoarse, qfine, aux_coarse, aux_fine, igrid) &
            bind(C, name="gfref_average2coarse")
        integer(c_int), value :: mx, my, mbc, maux, igrid
        real(c_double) :: qcoarse(*), qfine(*), aux_coarse(*), aux_fine(*)
        call fc2d_geoclaw_fort_average2coarse(mx, my, mbc, 3, qcoarse, qfine, maux, aux_coarse, &
            aux_fine, 0, igrid)
    end subroutine gfref_average2coarse

    integer(c_int) function gfref_timeinterp(mx, my, mbc, meqn, psize, qcurr, qlast, qinterp, alpha) &
            bind(C, name="gfref_timeinterp")
        integer(c_int), value :: mx, my, mbc, meqn, psize
        real(c_double) :: qcurr(*), qlast(*), qinterp(*)
        real(c_double), value :: alpha
        integer :: ierror
        call fc2d_geoclaw_fort_timeinterp(mx, my, mbc, meqn, psize, qcurr, qlast, qinterp, alpha, ierror)
        gfref_timeinterp = ierror
    end function gfref_timeinterp

end module gfref_amr_driver
