! ref_driver.f90 -- C-callable entry points into the reference GeoClaw
! Fortran step, compiled from the reference sources by oracle/Makefile into
! oracle/_ref/libgfref.so.
!
! TEST INFRASTRUCTURE ONLY: used to pin the C restatement (gf_oracle.c) and
! as the "reference" CPU baseline.  Module variables that the reference reads
! from geoclaw.data / friction data files (geoclaw_set_modules_fort.f90:1-63,
! geoclaw_module.f90:57-161) are set directly by gfref_set_params.

module gfref_driver
    use iso_c_binding
    implicit none
    integer, save :: gfref_src_term = 1
    external :: fc2d_geoclaw_rpn2, fc2d_geoclaw_rpt2
contains

    subroutine gfref_set_params(grav_in, drytol_in, manning_in, fdepth_in, &
            friction_in, src_term_in, order_in, transverse_in, mthlim_in) &
            bind(C, name="gfref_set_params")
        use geoclaw_module, only: grav, dry_tolerance, manning_coefficient, &
            manning_break, num_manning, friction_depth, friction_forcing, &
            coriolis_forcing, coordinate_system, rho, earth_radius
        use amr_module, only: mwaves, mcapa, method, mthlim, use_fwaves, rinfinity
        use friction_module, only: variable_friction
        real(c_double), value :: grav_in, drytol_in, manning_in, fdepth_in
        integer(c_int), value :: friction_in, src_term_in, order_in, transverse_in
        integer(c_int), intent(in) :: mthlim_in(3)

        grav = grav_in
        dry_tolerance = drytol_in
        if (allocated(manning_coefficient)) deallocate(manning_coefficient)
        if (allocated(manning_break)) deallocate(manning_break)
        num_manning = 1
        allocate(manning_coefficient(1), manning_break(1))
        manning_coefficient(1) = manning_in
        manning_break(1) = rinfinity
        friction_depth = fdepth_in
        friction_forcing = (friction_in /= 0)
        coriolis_forcing = .false.
        coordinate_system = 1
        earth_radius = 6367.5d3
        if (.not. allocated(rho)) allocate(rho(1))
        rho(1) = 1025.d0
        mwaves = 3
        mcapa = 0
        method = 0
        method(2) = order_in
        method(3) = transverse_in
        use_fwaves = .false.
        if (allocated(mthlim)) deallocate(mthlim)
        allocate(mthlim(3))
        mthlim(1:3) = mthlim_in(1:3)
        variable_friction = .false.
        gfref_src_term = src_term_in
    end subroutine gfref_set_params

    subroutine gfref_riemanntype(hL, hR, uL, uR, maxiter, drytol, g, out, rare) &
            bind(C, name="gfref_riemanntype")
        real(c_double), value :: hL, hR, uL, uR, drytol, g
        integer(c_int), value :: maxiter
        real(c_double), intent(out) :: out(3)
        integer(c_int), intent(out) :: rare(2)
        logical :: rare1, rare2
        integer :: mi
        double precision :: hm, s1m, s2m
        mi = maxiter
        call riemanntype(hL, hR, uL, uR, hm, s1m, s2m, rare1, rare2, mi, drytol, g)
        out(1) = hm
        out(2) = s1m
        out(3) = s2m
        rare(1) = merge(1, 0, rare1)
        rare(2) = merge(1, 0, rare2)
    end subroutine gfref_riemanntype

    subroutine gfref_riemann_aug_jcp(maxiter, vin, drytol, g, sw, fw) &
            bind(C, name="gfref_riemann_aug_jcp")
        integer(c_int), value :: maxiter
        real(c_double), intent(in) :: vin(18)
        real(c_double), value :: drytol, g
        real(c_double), intent(out) :: sw(3), fw(3, 3)
        double precision :: rho(1), sE1, sE2
        integer :: mi
        mi = maxiter
        rho(1) = 1025.d0
        sE1 = vin(17)
        sE2 = vin(18)
        call riemann_aug_JCP(mi, 3, 3, vin(1), vin(2), vin(3), vin(4), vin(5), &
            vin(6), vin(7), vin(8), vin(9), vin(10), vin(11), vin(12), vin(13), &
            vin(14), vin(15), vin(16), sE1, sE2, drytol, g, rho, sw, fw, 0, 0, 1)
    end subroutine gfref_riemann_aug_jcp

    subroutine gfref_rpn2(ixy, mx, mbc, maux, q1d, aux2, fwave, s, amdq, apdq) &
            bind(C, name="gfref_rpn2")
        integer(c_int), value :: ixy, mx, mbc, maux
        real(c_double) :: q1d(*), aux2(*), fwave(*), s(*), amdq(*), apdq(*)
        call fc2d_geoclaw_rpn2(ixy, mx, 3, 3, maux, mbc, mx, q1d, q1d, &
            aux2, aux2, fwave, s, amdq, apdq)
    end subroutine gfref_rpn2

    subroutine gfref_rpt2(ixy, imp, mx, mbc, maux, q1d, aux1, aux2, aux3, &
            asdq, bmasdq, bpasdq) bind(C, name="gfref_rpt2")
        integer(c_int), value :: ixy, imp, mx, mbc, maux
        real(c_double) :: q1d(*), aux1(*), aux2(*), aux3(*), asdq(*)
        real(c_double) :: bmasdq(*), bpasdq(*)
        call fc2d_geoclaw_rpt2(ixy, imp, mx, 3, 3, maux, mbc, mx, q1d, q1d, &
            aux1, aux2, aux3, asdq, bmasdq, bpasdq)
    end subroutine gfref_rpt2

    real(c_double) function gfref_philim(a, b, meth) bind(C, name="gfref_philim")
        real(c_double), value :: a, b
        integer(c_int), value :: meth
        double precision, external :: philim
        integer :: mm
        mm = meth
        gfref_philim = philim(a, b, mm)
    end function gfref_philim

    subroutine gfref_limiter(mx, mbc, wave, s, mthlim_in) bind(C, name="gfref_limiter")
        integer(c_int), value :: mx, mbc
        real(c_double) :: wave(*), s(*)
        integer(c_int), intent(in) :: mthlim_in(3)
        integer :: ml(3)
        ml = mthlim_in
        call fc2d_geoclaw_limiter(mx, 3, 3, mbc, mx, wave, s, ml)
    end subroutine gfref_limiter

    subroutine gfref_flux2(ixy, mx, mbc, maux, q1d, dtdx1d, aux1, aux2, aux3, &
            faddm, faddp, gaddm, gaddp, cfl1d) bind(C, name="gfref_flux2")
        integer(c_int), value :: ixy, mx, mbc, maux
        real(c_double) :: q1d(*), dtdx1d(*), aux1(*), aux2(*), aux3(*)
        real(c_double) :: faddm(*), faddp(*), gaddm(*), gaddp(*)
        real(c_double), intent(out) :: cfl1d
        call fc2d_geoclaw_flux2(ixy, mx, 3, maux, mbc, mx, q1d, dtdx1d, &
            aux1, aux2, aux3, faddm, faddp, gaddm, gaddp, cfl1d, &
            fc2d_geoclaw_rpn2, fc2d_geoclaw_rpt2)
    end subroutine gfref_flux2

    real(c_double) function gfref_step2_wrap(mx, my, mbc, maux, qold, aux, dx, dy, dt) &
            bind(C, name="gfref_step2_wrap")
        use amr_module, only: method, mthlim, mcapa, mwaves
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double) :: qold(*), aux(*)
        real(c_double), value :: dx, dy, dt
        double precision :: fp(3*(mx+2*mbc)*(my+2*mbc)), fm(3*(mx+2*mbc)*(my+2*mbc))
        double precision :: gp(3*(mx+2*mbc)*(my+2*mbc)), gm(3*(mx+2*mbc)*(my+2*mbc))
        double precision :: work(1), cfl, xlower, ylower, t
        integer :: maxm, mwork, level, block_corner_count(0:3), mthl(3), meth(7)
        maxm = max(mx, my)
        mwork = 1
        level = 0
        block_corner_count = 0
        xlower = 0.d0
        ylower = 0.d0
        t = 0.d0
        cfl = 0.d0
        mthl = mthlim(1:3)
        meth = method
        call fc2d_geoclaw_step2_wrap(maxm, 3, maux, mbc, meth, mthl, mcapa, &
            mwaves, mx, my, qold, aux, dx, dy, dt, cfl, work, mwork, xlower, &
            ylower, level, t, fp, fm, gp, gm, fc2d_geoclaw_rpn2, &
            fc2d_geoclaw_rpt2, block_corner_count)
        gfref_step2_wrap = cfl
    end function gfref_step2_wrap

    subroutine gfref_b4step2(mx, my, mbc, maux, q, aux, dt) bind(C, name="gfref_b4step2")
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double) :: q(*), aux(*)
        real(c_double), value :: dt
        integer :: lmx, lmy, lmbc, lmaux
        double precision :: xl, yl, dxx, dyy, t, ldt
        lmx = mx
        lmy = my
        lmbc = mbc
        lmaux = maux
        xl = 0.d0
        yl = 0.d0
        dxx = 1.d0
        dyy = 1.d0
        t = 0.d0
        ldt = dt
        call fc2d_geoclaw_b4step2(lmbc, lmx, lmy, 3, q, xl, yl, dxx, dyy, t, ldt, lmaux, aux)
    end subroutine gfref_b4step2

    subroutine gfref_src2(mx, my, mbc, maux, q, aux, dt) bind(C, name="gfref_src2")
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double) :: q(*), aux(*)
        real(c_double), value :: dt
        call fc2d_geoclaw_src2(3, mbc, mx, my, 0.d0, 0.d0, 1.d0, 1.d0, q, maux, aux, 0.d0, dt)
    end subroutine gfref_src2

    subroutine gfref_bc2(mx, my, mbc, maux, q, aux, mthbc) bind(C, name="gfref_bc2")
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double) :: q(*), aux(*)
        integer(c_int), intent(in) :: mthbc(4)
        integer :: mb(4)
        mb = mthbc
        call fc2d_geoclaw_bc2(3, mbc, mx, my, 0.d0, 0.d0, 1.d0, 1.d0, q, maux, aux, &
            0.d0, 0.d0, mb)
    end subroutine gfref_bc2

    ! Hydrograph module state that read_hydrograph_data would read from
    ! hydrograph.data (hydrograh_module.f90:35-144): discharge type, the
    ! table, the channel strip and the enabled sides (left, right, bottom,
    ! top -> boundary_location(1), (2), (4), (3)).
    subroutine gfref_set_hydrograph(nrows, t_in, hu_in, b_in, x0_in, y0_in, sides) &
            bind(C, name="gfref_set_hydrograph")
        use hydrograph_module, only: use_hydrograph, hydrograph_type, boundary_location, &
            hb => b, hx0 => x0, hy0 => y0, htime => time, heta => eta, hhu => hu, num_rows
        integer(c_int), value :: nrows
        real(c_double), intent(in) :: t_in(nrows), hu_in(nrows)
        real(c_double), value :: b_in, x0_in, y0_in
        integer(c_int), intent(in) :: sides(4)
        use_hydrograph = 'True'
        hydrograph_type = 'discharge'
        boundary_location(:) = 'none'
        if (sides(1) /= 0) boundary_location(1) = 'left'
        if (sides(2) /= 0) boundary_location(2) = 'right'
        if (sides(3) /= 0) boundary_location(4) = 'bottom'
        if (sides(4) /= 0) boundary_location(3) = 'top'
        hb = b_in
        hx0 = x0_in
        hy0 = y0_in
        num_rows = nrows
        if (allocated(htime)) deallocate(htime)
        if (allocated(heta)) deallocate(heta)
        if (allocated(hhu)) deallocate(hhu)
        allocate(htime(nrows), heta(nrows), hhu(nrows))
        htime = t_in
        hhu = hu_in
        heta = 0.d0
    end subroutine gfref_set_hydrograph

    ! inflow_interpolate (hydrograh_module.f90:155-168): returns q0(2).
    real(c_double) function gfref_hydro_interpolate(t) bind(C, name="gfref_hydro_interpolate")
        use hydrograph_module, only: inflow_interpolate
        real(c_double), value :: t
        double precision :: q0(4), q1(4)
        q0 = 0.d0
        q1 = 0.d0
        call inflow_interpolate(t, q0, q1)
        gfref_hydro_interpolate = q0(2)
    end function gfref_hydro_interpolate

    subroutine gfref_riemann_invariants(q0, q1) bind(C, name="gfref_riemann_invariants")
        use hydrograph_module, only: Riemann_invariants
        real(c_double), intent(inout) :: q0(2)
        real(c_double), intent(in) :: q1(2)
        double precision :: a(4), c(4)
        a = 0.d0
        c = 0.d0
        a(1:2) = q0
        c(1:2) = q1
        call Riemann_invariants(a, c)
        q0 = a(1:2)
    end subroutine gfref_riemann_invariants

    subroutine gfref_two_shock(q0, q1) bind(C, name="gfref_two_shock")
        use hydrograph_module, only: two_shock
        real(c_double), intent(inout) :: q0(2)
        real(c_double), intent(in) :: q1(2)
        double precision :: a(4), c(4)
        a = 0.d0
        c = 0.d0
        a(1:2) = q0
        c(1:2) = q1
        call two_shock(a, c)
        q0 = a(1:2)
    end subroutine gfref_two_shock

    ! fc2d_geoclaw_bc2 with the patch geometry and time (mthbc = 0 uses the
    ! hydrograph set above).
    subroutine gfref_bc2_ex(mx, my, mbc, maux, q, aux, mthbc, xlower, ylower, dx, dy, t) &
            bind(C, name="gfref_bc2_ex")
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double) :: q(*), aux(*)
        integer(c_int), intent(in) :: mthbc(4)
        real(c_double), value :: xlower, ylower, dx, dy, t
        integer :: mb(4)
        mb = mthbc
        call fc2d_geoclaw_bc2(3, mbc, mx, my, xlower, ylower, dx, dy, q, maux, aux, &
            t, 0.d0, mb)
    end subroutine gfref_bc2_ex

    ! Refinement criteria that read_refinement_data / set_flow_grades read
    ! from refinement.data / flowgrades.data (refinement_data_module.f90),
    ! the sea level, and the regions common block (fclaw2d_source/regions.i).
    subroutine gfref_set_refinement(wave_tol, deep, max_lev_deep, max_vdp, nspeeds, speeds, &
            nfg, fgvalue, fgvar, fgminlevel, sealevel, nreg, reg) bind(C, name="gfref_set_refinement")
        use refinement_data_module, only: wave_tolerance, deep_depth, max_level_deep, &
            max_velocity_depth_product, speed_tolerance, num_flowgrades, flowgradevalue, &
            iflowgradevariable, iflowgradetype, iflowgrademinlevel
        use geoclaw_module, only: sea_level
        real(c_double), value :: wave_tol, deep, max_vdp, sealevel
        integer(c_int), value :: max_lev_deep, nspeeds, nfg, nreg
        real(c_double), intent(in) :: speeds(nspeeds), fgvalue(*), reg(8, *)
        integer(c_int), intent(in) :: fgvar(*), fgminlevel(*)
        integer :: m
        include 'regions.i'
        wave_tolerance = wave_tol
        deep_depth = deep
        max_level_deep = max_lev_deep
        max_velocity_depth_product = max_vdp
        if (allocated(speed_tolerance)) deallocate(speed_tolerance)
        allocate(speed_tolerance(nspeeds))
        speed_tolerance = speeds
        num_flowgrades = nfg
        if (allocated(flowgradevalue)) deallocate(flowgradevalue)
        if (allocated(iflowgradevariable)) deallocate(iflowgradevariable)
        if (allocated(iflowgradetype)) deallocate(iflowgradetype)
        if (allocated(iflowgrademinlevel)) deallocate(iflowgrademinlevel)
        allocate(flowgradevalue(max(nfg, 1)), iflowgradevariable(max(nfg, 1)), &
            iflowgradetype(max(nfg, 1)), iflowgrademinlevel(max(nfg, 1)))
        do m = 1, nfg
            flowgradevalue(m) = fgvalue(m)
            iflowgradevariable(m) = fgvar(m)
            iflowgradetype(m) = 1
            iflowgrademinlevel(m) = fgminlevel(m)
        end do
        sea_level = sealevel
        mregions = nreg
        do m = 1, nreg
            xlowregion(m) = reg(1, m)
            ylowregion(m) = reg(2, m)
            xhiregion(m) = reg(3, m)
            yhiregion(m) = reg(4, m)
            tlowregion(m) = reg(5, m)
            thiregion(m) = reg(6, m)
            minlevelregion(m) = int(reg(7, m))
            maxlevelregion(m) = int(reg(8, m))
        end do
    end subroutine gfref_set_refinement

    integer(c_int) function gfref_tag4refinement(mx, my, mbc, maux, xlower, ylower, dx, dy, t, &
            q, aux, level, maxlevel, init_flag) bind(C, name="gfref_tag4refinement")
        integer(c_int), value :: mx, my, mbc, maux, level, maxlevel, init_flag
        real(c_double), value :: xlower, ylower, dx, dy, t
        real(c_double) :: q(*), aux(*)
        integer :: tag
        call fc2d_geoclaw_fort_tag4refinement(mx, my, mbc, 3, maux, xlower, ylower, dx, dy, t, &
            0, q, aux, level, maxlevel, init_flag, tag)
        gfref_tag4refinement = tag
    end function gfref_tag4refinement

    integer(c_int) function gfref_tag4coarsening(mx, my, mbc, maux, xlower, ylower, dx, dy, t, &
            q0, q1, q2, q3, a0, a1, a2, a3, level, maxlevel, init_flag) &
            bind(C, name="gfref_tag4coarsening")
        integer(c_int), value :: mx, my, mbc, maux, level, maxlevel, init_flag
        real(c_double), intent(in) :: xlower(4), ylower(4)
        real(c_double), value :: dx, dy, t
        real(c_double) :: q0(*), q1(*), q2(*), q3(*), a0(*), a1(*), a2(*), a3(*)
        integer :: tag
        double precision :: xl(0:3), yl(0:3)
        xl = xlower
        yl = ylower
        call fc2d_geoclaw_fort_tag4coarsening(0, mx, my, mbc, 3, maux, xl, yl, dx, dy, t, &
            q0, q1, q2, q3, a0, a1, a2, a3, level, maxlevel, init_flag, tag)
        gfref_tag4coarsening = tag
    end function gfref_tag4coarsening

    ! ASCII frame output (fc2d_geoclaw_output_ascii_fort.f): the header routine
    ! writes fort.tNNNN and empties fort.qNNNN, the patch routine appends one
    ! patch to fort.qNNNN, both in the process's working directory.  The patch
    ! routine zeroes |q| < 1e-99 in q itself.
    subroutine gfref_write_header(iframe, time, meqn, maux, ngrids, mbc) &
            bind(C, name="gfref_write_header")
        integer(c_int), value :: iframe, meqn, maux, ngrids, mbc
        real(c_double), value :: time
        call fc2d_geoclaw_fort_write_header(iframe, time, meqn, maux, ngrids, mbc)
    end subroutine gfref_write_header

    subroutine gfref_write_file(mx, my, meqn, maux, mbc, xlower, ylower, dx, dy, q, aux, &
            iframe, patch_num, level, blockno, mpirank) bind(C, name="gfref_write_file")
        integer(c_int), value :: mx, my, meqn, maux, mbc, iframe, patch_num, level, blockno, mpirank
        real(c_double), value :: xlower, ylower, dx, dy
        real(c_double) :: q(*), aux(*)
        call fc2d_geoclaw_fort_write_file(mx, my, meqn, maux, mbc, xlower, ylower, dx, dy, &
            q, aux, iframe, patch_num, level, blockno, mpirank)
    end subroutine gfref_write_file

    ! Gauge sample of one patch (geoclaw_gauges_fort.f90): q and bathymetry at
    ! (xc, yc), bilinear unless a neighbouring depth is below dry_tolerance/10.
    subroutine gfref_update_gauge(mx, my, mbc, maux, xlower, ylower, dx, dy, q, aux, &
            xc, yc, qvar, avar) bind(C, name="gfref_update_gauge")
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double), value :: xlower, ylower, dx, dy, xc, yc
        real(c_double) :: q(*), aux(*), qvar(3), avar(*)
        call fc2d_geoclaw_update_gauge(mx, my, mbc, 3, xlower, ylower, dx, dy, q, maux, aux, &
            xc, yc, qvar, avar)
    end subroutine gfref_update_gauge

    ! geoclaw_update (fc2d_cpucuda.cpp:362-406): b4step2, step2, MAXCFL_CAP,
    ! src2.  Signature matches gfo_update_fn so the C brick driver can run it.
    real(c_double) function gfref_patch_update(ctx, mx, my, mbc, maux, q, aux, &
            dx, dy, t, dt) bind(C, name="gfref_patch_update")
        type(c_ptr), value :: ctx
        integer(c_int), value :: mx, my, mbc, maux
        real(c_double) :: q(*), aux(*)
        real(c_double), value :: dx, dy, t, dt
        double precision :: maxcfl
        call gfref_b4step2(mx, my, mbc, maux, q, aux, dt)
        maxcfl = gfref_step2_wrap(mx, my, mbc, maux, q, aux, dx, dy, dt)
        if (maxcfl > 1000.d0) maxcfl = 0.d0
        if (gfref_src_term > 0) call gfref_src2(mx, my, mbc, maux, q, aux, dt)
        gfref_patch_update = maxcfl
    end function gfref_patch_update

end module gfref_driver
