!> Drop-in replacements for the reference's hot-path modules, with the
!> reference's module names, public lists and argument lists, running on the
!> GPU through libdcm_hip.so.  A model build links this file (and dcm_hip.f90)
!> in place of pamm/sfvt/advection_3d.F90, dynamics.f90 and diagnostics.f90.
!>
!> One process, one GPU, one MPI rank: dims must be (1,1,1); comm3d, id, rank
!> and the communicators are accepted and unused.  The halo is one cell
!> (l_h = r_h = 1, main.f90:80-81).  Arrays stay owned by the caller; each
!> call uploads its 3-D inputs into device buffers kept per argument slot,
!> runs the kernels, and copies its outputs back.  Options whose physics is
!> not on the GPU path (viscous, damping layer, forcing, divergence
!> forcing, radiation, the vertical-only MPDATA of the microphysics) stop
!> with a message, as the reference stops on its own errors.

module dcm_dropin_util
    use iso_c_binding
    use dcm_hip
    implicit none
    private
    public :: one_rank, to_cint, ones

contains

    !> the GPU path is single-rank
    subroutine one_rank(dims)
        integer(dcm_i4b), intent(in) :: dims(3)
        if (dims(1) /= 1 .or. dims(2) /= 1 .or. dims(3) /= 1) then
            print *, 'dcm: the GPU path runs one rank per GPU (dims must be 1,1,1)'
            error stop 1
        endif
    end subroutine one_rank

    integer(c_int) function to_cint(flag)
        logical, intent(in) :: flag
        to_cint = 0
        if (flag) to_cint = 1
    end function to_cint

    !> a metric stand-in of ones, lower bound 0
    function ones(n) result(a)
        integer, intent(in) :: n
        real(c_double) :: a(0:n+1)
        a = 1.0_c_double
    end function ones

end module dcm_dropin_util


!> module advection_s_3d (pamm/sfvt/advection_3d.F90:5-18)
module advection_s_3d
    use iso_c_binding
    use dcm_hip
    use dcm_dropin_util
    implicit none
    private
    public :: mpdata_3d, mpdata_vec_3d, first_order_upstream_3d, adv_ref_state, &
        mpdata_3d_add, mpdata_vert_3d, mpdata_vec_vert_3d

contains

    !> first_order_upstream_3d (pamm/sfvt/advection_3d.F90:39-116)
    subroutine first_order_upstream_3d(dt, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, &
            u, v, w, psi, neumann, dims, coords)
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(in), target :: u
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: v
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: psi
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dzn, rhoa, rhoan
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        integer(dcm_i4b), intent(in) :: neumann
        type(c_ptr) :: ctx, du, dv, dw, dpsi
        integer(c_size_t) :: nb

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ! the donor cell reads dxn, dyn, dzn, rhoa, rhoan only
        ctx = dcm_context(ip, jp, kp, dxn, dyn, dzn, dxn, dyn, dzn, rhoa, rhoan, &
            provisional=.true.)
        nb = field_bytes(ip, jp, kp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 4, c_loc(psi), nb)
        call dcm_check(dcm_upwind3d(ctx, dt, du, dv, dw, dpsi, int(neumann, c_int)), &
            'first_order_upstream_3d')
        call dcm_take(ctx, 4, c_loc(psi), nb)
    end subroutine first_order_upstream_3d

    !> mpdata_3d (pamm/sfvt/advection_3d.F90:219-714)
    subroutine mpdata_3d(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, &
            u, v, w, psi_in, lbc, ubc, kord, monotone, neumann, comm3d, id, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h, kord
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(in), target :: u
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: v
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: psi_in
        real(dcm_wp), intent(inout) :: lbc, ubc
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dx, dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        logical, intent(in) :: monotone
        integer(dcm_i4b), intent(in) :: neumann
        type(c_ptr) :: ctx, du, dv, dw, dpsi
        integer(c_size_t) :: nb
        integer(c_int) :: early

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ctx = dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ip, jp, kp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 4, c_loc(psi_in), nb)
        call dcm_check(dcm_mpdata3d(ctx, dt, du, dv, dw, dpsi, lbc, ubc, int(kord, c_int), &
            to_cint(monotone), int(neumann, c_int), early), 'mpdata_3d')
        call dcm_take(ctx, 4, c_loc(psi_in), nb)
    end subroutine mpdata_3d

    !> mpdata_vec_3d (pamm/sfvt/advection_3d.F90:1250-1753): nq fields of one
    !> category, advected with field 1's corrected velocities
    subroutine mpdata_vec_3d(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, nq, l_h, &
            r_h, u, v, w, psi_in, lbc, ubc, kord, monotone, neumann, comm3d, id, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, nq, l_h, r_h, kord
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(in), target :: u
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: v
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h, 1:nq), &
            intent(inout), target :: psi_in
        real(dcm_wp), intent(inout), dimension(nq) :: lbc, ubc
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dx, dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        logical, intent(in) :: monotone
        integer(dcm_i4b), intent(in) :: neumann
        type(c_ptr) :: ctx, du, dv, dw, dpsi
        integer(c_size_t) :: nb
        integer(c_int) :: early

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        if (nq < 1) return
        ctx = dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ip, jp, kp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 16, c_loc(psi_in), nb * nq)
        call dcm_check(dcm_mpdata_vec3d(ctx, dt, du, dv, dw, dpsi, int(nq, c_int), lbc, ubc, &
            int(kord, c_int), to_cint(monotone), int(neumann, c_int), early), 'mpdata_vec_3d')
        call dcm_take(ctx, 16, c_loc(psi_in), nb * nq)
    end subroutine mpdata_vec_3d

    !> adv_ref_state (pamm/sfvt/advection_3d.F90:1148-1218).  dy and dyn are
    !> declared with the ip extent there (:1158) and are not read.
    subroutine adv_ref_state(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, &
            u, v, w, psi_3d, psi_ref, comm3d, id, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dx, dxn
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: rhoa, rhoan, dz, dzn, psi_ref
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: u, v, w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: psi_3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        type(c_ptr) :: ctx, dw, dpsi
        integer(c_size_t) :: nb

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ! the stencil reads dz, rhoa and rhoan only; dy, dyn have the wrong extent
        ctx = dcm_context(ip, jp, kp, dx, ones(jp), dz, dxn, ones(jp), dzn, rhoa, rhoan, &
            provisional=.true.)
        nb = field_bytes(ip, jp, kp)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 4, c_loc(psi_3d), nb)
        call dcm_check(dcm_adv_ref_state(ctx, dt, dw, dpsi, psi_ref), 'adv_ref_state')
        call dcm_take(ctx, 4, c_loc(psi_3d), nb)
    end subroutine adv_ref_state

    !> mpdata_3d_add (pamm/sfvt/advection_3d.F90:139-190): psi_1d added to
    !> every column (halo levels included), mpdata_3d, psi_1d removed
    subroutine mpdata_3d_add(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, &
            u, v, w, psi_in, psi_1d, lbc, ubc, kord, monotone, neumann, comm3d, id, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h, kord
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(in), target :: u
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: v
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: psi_in
        real(dcm_wp), dimension(-r_h+1:kp+r_h), intent(in) :: psi_1d
        real(dcm_wp), intent(inout) :: lbc, ubc
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dx, dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        logical, intent(in) :: monotone
        integer(dcm_i4b), intent(in) :: neumann
        type(c_ptr) :: ctx, du, dv, dw, dpsi
        integer(c_size_t) :: nb
        integer(c_int) :: early

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ctx = dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ip, jp, kp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 4, c_loc(psi_in), nb)
        call dcm_check(dcm_mpdata3d_add(ctx, dt, du, dv, dw, dpsi, psi_1d, lbc, ubc, &
            int(kord, c_int), to_cint(monotone), int(neumann, c_int), early), 'mpdata_3d_add')
        call dcm_take(ctx, 4, c_loc(psi_in), nb)
    end subroutine mpdata_3d_add

    !> mpdata_vert_3d (pamm/sfvt/advection_3d.F90:820-1126): vertical-only
    !> MPDATA, the sedimentation transport of the microphysics
    subroutine mpdata_vert_3d(dt, dz, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, w, psi_in, lbc, &
            ubc, kord, monotone, neumann, comm3d, id, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h, kord
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: psi_in
        real(dcm_wp), intent(inout) :: lbc, ubc
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        logical, intent(in) :: monotone
        integer(dcm_i4b), intent(in) :: neumann
        type(c_ptr) :: ctx, dw, dpsi
        integer(c_size_t) :: nb
        integer(c_int) :: early

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ! only the vertical metrics arrive here: a provisional context
        ctx = dcm_context(ip, jp, kp, ones(ip), ones(jp), dz, ones(ip), ones(jp), dzn, rhoa, &
            rhoan, provisional=.true.)
        nb = field_bytes(ip, jp, kp)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 4, c_loc(psi_in), nb)
        call dcm_check(dcm_mpdata_vert3d(ctx, dt, dw, dpsi, lbc, ubc, int(kord, c_int), &
            to_cint(monotone), int(neumann, c_int), early), 'mpdata_vert_3d')
        call dcm_take(ctx, 4, c_loc(psi_in), nb)
    end subroutine mpdata_vert_3d

    !> mpdata_vec_vert_3d (pamm/sfvt/advection_3d.F90:1779-2087): nq fields of
    !> one category falling with field 1's corrected velocities (the call of
    !> pamm/microphysics.f90:1522-1560)
    subroutine mpdata_vec_vert_3d(dt, dz, dzn, rhoa, rhoan, ip, jp, kp, nq, l_h, r_h, w, &
            psi_in, lbc, ubc, kord, monotone, neumann, comm3d, id, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, nq, l_h, r_h, kord
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h, 1:nq), &
            intent(inout), target :: psi_in
        real(dcm_wp), intent(inout), dimension(nq) :: lbc, ubc
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        logical, intent(in) :: monotone
        integer(dcm_i4b), intent(in) :: neumann
        type(c_ptr) :: ctx, dw, dpsi
        integer(c_size_t) :: nb
        integer(c_int) :: early

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        if (nq < 1) return
        ctx = dcm_context(ip, jp, kp, ones(ip), ones(jp), dz, ones(ip), ones(jp), dzn, rhoa, &
            rhoan, provisional=.true.)
        nb = field_bytes(ip, jp, kp)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dpsi = dcm_put(ctx, 16, c_loc(psi_in), nb * nq)
        call dcm_check(dcm_mpdata_vec_vert3d(ctx, dt, dw, dpsi, int(nq, c_int), lbc, ubc, &
            int(kord, c_int), to_cint(monotone), int(neumann, c_int), early), &
            'mpdata_vec_vert_3d')
        call dcm_take(ctx, 16, c_loc(psi_in), nb * nq)
    end subroutine mpdata_vec_vert_3d

end module advection_s_3d


!> module d_solver (dynamics.f90:5-21)
module d_solver
    use iso_c_binding
    use, intrinsic :: ieee_arithmetic, only: ieee_is_finite
    use dcm_hip
    use dcm_dropin_util
    implicit none
    private
    public :: bicgstab, sources, advance_momentum, radiative_transfer

contains

    !> bicgstab (dynamics.f90:293-444): x is the pressure (warm start), b the
    !> right-hand side; both are written back with their halos as the
    !> reference leaves them
    subroutine bicgstab(comm3d, id, rank, dims, coords, dt, xg, yg, zg, dx, dy, dz, dxn, dyn, &
            dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, su, sv, sw, zu, zv, zw, x, b, tol, &
            test_solver)
        integer(dcm_i4b), intent(in) :: id, comm3d, rank
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(in) :: su, zu
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in) :: sv, zv
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: sw, zw
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: x
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: b
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: xg, dx, dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: yg, dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: zg, dz, dzn, rhoa, rhoan
        logical :: test_solver
        real(dcm_wp), intent(in) :: tol
        type(c_ptr) :: ctx, dsw, dx3, db
        integer(c_size_t) :: nb
        integer(c_int) :: iters, st
        real(c_double) :: err

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ctx = dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ip, jp, kp)
        dsw = dcm_put(ctx, 10, c_loc(sw), nb)
        dx3 = dcm_put(ctx, 14, c_loc(x), nb)
        db = dcm_put(ctx, 15, c_loc(b), nb)
        st = dcm_bicgstab(ctx, xg, yg, zg, dsw, dx3, db, tol, to_cint(test_solver), 0_c_int, &
            iters, err)
        ! the reference prints and stops when itmax is reached (:431-434)
        if (st == DCM_NOT_CONVERGED) then
            print *, 'convergence failure with error,it ', err, iters
            stop
        endif
        if (st == DCM_NONFINITE) then
            ! name the input that carries the non-finite values
            call dcm_take(ctx, 14, c_loc(x), nb)
            call dcm_take(ctx, 15, c_loc(b), nb)
            print *, 'bicgstab: non-finite entries in the interior of x: ', &
                count(.not. ieee_is_finite(x(1:kp, 1:jp, 1:ip))), ', of b: ', &
                count(.not. ieee_is_finite(b(1:kp, 1:jp, 1:ip))), ', after ', iters, &
                ' iterations'
        endif
        call dcm_check(st, 'bicgstab')
        call dcm_take(ctx, 14, c_loc(x), nb)
        call dcm_take(ctx, 15, c_loc(b), nb)
    end subroutine bicgstab

    !> sources (dynamics.f90:639-1000): momentum advection, buoyancy, the
    !> damping layer, forcing, subsidence and the subgrid model
    !> (calculate_subgrid_3d) as the flags ask, halo fills, boundary rows and
    !> the Poisson right-hand side.  Radiation is not on the GPU path.
    subroutine sources(comm3d, id, rank, dims, coords, dt, xg, yg, zg, zng, dx, dy, dz, dxn, &
            dyn, dzn, ip, jp, kp, l_h, r_h, nq, ubar, vbar, wbar, thbar, qbar, dampfacn, &
            dampfac, u_force, v_force, forcing_tau, w_subs, zu, zv, zw, u, v, w, su, sv, sw, &
            q, sq, viss, rhs, th, sth, strain, vism, vist, theta, thetan, tref, trefn, rhoa, &
            rhoan, lamsq, lamsqn, z0, z0th, viscous, moisture, damping_layer, forcing, &
            divergence, radiation, sub_vert_comm, time, tdstart, tdend, a, b, c, r, usol, &
            nbands, ns, nl, flux_u, flux_d, rad_power, lambda, lambda_low, lambda_high, &
            delta_lambda, nrwbin, niwbin, sflux_l, b_s_g, start_year, start_mon, start_day, &
            start_hour, start_min, start_sec, lat, lon, albedo, emiss, quad_flag, &
            asymmetry_water, nrad, ngs, lamgs, mugs, nprocv, mvrecv)
        integer(dcm_i4b), intent(in) :: id, comm3d, rank
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        logical, intent(in) :: viscous, moisture, damping_layer, forcing, divergence, &
            radiation
        real(dcm_wp), intent(in) :: dt, z0, z0th, forcing_tau
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h, nq
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(in), target :: u, zu
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: v, zv
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: w, zw
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: su, sv, sw
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: rhs, sth, strain, vism, vist
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h, 1:nq), &
            intent(in), target :: q
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h, 1:nq), &
            intent(inout), target :: sq, viss
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: xg, dx, dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: yg, dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: zng, dz, dzn, tref, trefn, &
            rhoa, rhoan, lamsqn, ubar, vbar, wbar, thbar
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in), target :: zg, theta, thetan, &
            lamsq, dampfacn, dampfac, u_force, v_force, w_subs
        real(dcm_wp), dimension(-l_h+1:kp+r_h, nq), intent(in) :: qbar
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: th
        integer(dcm_i4b), intent(in) :: sub_vert_comm
        real(dcm_wp), intent(in) :: time
        integer(dcm_i4b), intent(in) :: nbands, ns, nl, tdstart, tdend
        integer(dcm_i4b), intent(in) :: nprocv
        real(dcm_wp), dimension(nbands) :: lambda, b_s_g, lambda_low, lambda_high, &
            delta_lambda, nrwbin, niwbin, sflux_l
        real(dcm_wp), intent(in), dimension(nprocv) :: mvrecv
        real(dcm_wp), intent(inout), dimension(1:tdend) :: a, b, c, r, usol
        real(dcm_wp), dimension(1-r_h:kp+r_h, 1-r_h:jp+r_h, 1-r_h:ip+r_h, 1:nbands), &
            intent(inout) :: flux_d, flux_u
        real(dcm_wp), dimension(1-r_h:kp+r_h, 1-r_h:jp+r_h, 1-r_h:ip+r_h) :: rad_power
        real(dcm_wp), intent(in) :: lat, lon, albedo, emiss, asymmetry_water
        integer(dcm_i4b), intent(in) :: quad_flag, start_year, start_mon, start_day, &
            start_hour, start_min, start_sec, nrad
        real(dcm_wp), dimension(1-r_h:kp+r_h, 1-r_h:jp+r_h, 1-r_h:ip+r_h, 1:nrad), &
            intent(in) :: ngs, lamgs, mugs
        type(c_ptr) :: ctx, du, dv, dw, dth, dsu, dsv, dsw, dsth, drhs, dzu, dzv, dzw, dq, dsq, &
            dstrain, dvism, dvist, dviss
        type(dcm_physics) :: ph
        integer(c_size_t) :: nb
        integer(c_int) :: nqm

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        if (radiation) then
            print *, 'dcm: radiation is not on the GPU path'
            stop
        endif
        ctx = dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ip, jp, kp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dth = dcm_put(ctx, 13, c_loc(th), nb)
        ! inout: the cells sources does not write keep the caller's values
        dsu = dcm_put(ctx, 8, c_loc(su), nb)
        dsv = dcm_put(ctx, 9, c_loc(sv), nb)
        dsw = dcm_put(ctx, 10, c_loc(sw), nb)
        dsth = dcm_put(ctx, 11, c_loc(sth), nb)
        drhs = dcm_put(ctx, 12, c_loc(rhs), nb)
        if (.not. (viscous .or. damping_layer .or. forcing .or. divergence)) then
            call dcm_check(dcm_sources(ctx, du, dv, dw, dth, theta, thbar, dsu, dsv, dsw, dsth, &
                drhs), 'sources')
        else
            ph%viscous = to_cint(viscous)
            ph%damping_layer = to_cint(damping_layer)
            ph%forcing = to_cint(forcing)
            ph%divergence = to_cint(divergence)
            ph%forcing_tau = forcing_tau
            ph%z0 = z0
            ph%z0th = z0th
            ph%lamsq = c_loc(lamsq)
            ph%dampfacn = c_loc(dampfacn)
            ph%dampfac = c_loc(dampfac)
            ph%u_force = c_loc(u_force)
            ph%v_force = c_loc(v_force)
            ph%w_subs = c_loc(w_subs)
            ph%z = c_loc(zg)
            ph%theta = c_loc(theta)
            ph%thetan = c_loc(thetan)
            nqm = 0
            if (moisture) nqm = int(nq, c_int)
            dzu = dcm_put(ctx, 4, c_loc(zu), nb)
            dzv = dcm_put(ctx, 5, c_loc(zv), nb)
            dzw = dcm_put(ctx, 6, c_loc(zw), nb)
            dstrain = dcm_put(ctx, 19, c_loc(strain), nb)
            dvism = dcm_put(ctx, 20, c_loc(vism), nb)
            dvist = dcm_put(ctx, 21, c_loc(vist), nb)
            dq = c_null_ptr
            dsq = c_null_ptr
            dviss = c_null_ptr
            if (nqm > 0) then
                dq = dcm_put(ctx, 17, c_loc(q), nb * nqm)
                dsq = dcm_put(ctx, 18, c_loc(sq), nb * nqm)
                dviss = dcm_put(ctx, 22, c_loc(viss), nb * nqm)
            endif
            call dcm_check(dcm_sources_phys(ctx, ph, ubar, vbar, wbar, du, dv, dw, dzu, dzv, &
                dzw, dth, dq, nqm, theta, thbar, dsu, dsv, dsw, dsth, dsq, dstrain, dvism, &
                dvist, dviss, drhs), 'sources')
            if (viscous) then
                call dcm_take(ctx, 19, c_loc(strain), nb)
                call dcm_take(ctx, 20, c_loc(vism), nb)
                call dcm_take(ctx, 21, c_loc(vist), nb)
                if (nqm > 0) then
                    call dcm_take(ctx, 18, c_loc(sq), nb * nqm)
                    call dcm_take(ctx, 22, c_loc(viss), nb * nqm)
                endif
            endif
        endif
        call dcm_take(ctx, 8, c_loc(su), nb)
        call dcm_take(ctx, 9, c_loc(sv), nb)
        call dcm_take(ctx, 10, c_loc(sw), nb)
        call dcm_take(ctx, 11, c_loc(sth), nb)
        call dcm_take(ctx, 12, c_loc(rhs), nb)
    end subroutine sources

    !> advance_momentum (dynamics.f90:1019-1099); dt is the leapfrog step 2*dt
    !> the driver passes
    subroutine advance_momentum(comm3d, id, rank, dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, &
            ip, jp, kp, l_h, r_h, u, v, w, zu, zv, zw, su, sv, sw, p, dims, coords)
        integer(dcm_i4b), intent(in) :: id, comm3d, rank
        real(dcm_wp), intent(in) :: dt
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -l_h+1:ip+r_h), &
            intent(inout), target :: u, zu
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -l_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: v, zv
        real(dcm_wp), dimension(-l_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout), target :: w, zw
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(in), target :: su, sv, sw, p
        real(dcm_wp), dimension(-l_h+1:ip+r_h), intent(in) :: dx, dxn
        real(dcm_wp), dimension(-l_h+1:jp+r_h), intent(in) :: dy, dyn
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        type(c_ptr) :: ctx, du, dv, dw, dzu, dzv, dzw, dsu, dsv, dsw, dp
        integer(c_size_t) :: nb

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ctx = dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ip, jp, kp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dzu = dcm_put(ctx, 5, c_loc(zu), nb)
        dzv = dcm_put(ctx, 6, c_loc(zv), nb)
        dzw = dcm_put(ctx, 7, c_loc(zw), nb)
        dsu = dcm_put(ctx, 8, c_loc(su), nb)
        dsv = dcm_put(ctx, 9, c_loc(sv), nb)
        dsw = dcm_put(ctx, 10, c_loc(sw), nb)
        dp = dcm_put(ctx, 17, c_loc(p), nb)
        call dcm_check(dcm_advance_momentum(ctx, dt, du, dv, dw, dzu, dzv, dzw, dsu, dsv, dsw, &
            dp), 'advance_momentum')
        call dcm_take(ctx, 1, c_loc(u), nb)
        call dcm_take(ctx, 2, c_loc(v), nb)
        call dcm_take(ctx, 3, c_loc(w), nb)
    end subroutine advance_momentum

    !> radiative_transfer (dynamics.f90:491-630) is not on the GPU path
    subroutine radiative_transfer(comm3d, id, rank, dims, coords, dt, dz, dzn, ip, jp, kp, &
            l_h, r_h, th, sth, tref, trefn, rhoa, rhoan, sub_vert_comm, time, tdstart, tdend, &
            a, b, c, r, usol, nbands, ns, nl, flux_u, flux_d, lambda, lambda_low, &
            lambda_high, delta_lambda, nrwbin, niwbin, sflux_l, b_s_g, nq, q, start_year, &
            start_mon, start_day, start_hour, start_min, start_sec, lat, lon, albedo, emiss, &
            quad_flag, gas_absorption, nmolecule, nweights, npress, ntemp, nh2o, itemp, &
            ipress, bli_read, probs_read, h2o_read, moleculeID, moleculePPM, &
            molecularWeights, asymmetry_water, nrad, ngs, lamgs, mugs, nprocv, mvrecv)
        integer(dcm_i4b), intent(in) :: id, comm3d, rank
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        real(dcm_wp), intent(in) :: dt
        logical, intent(in) :: gas_absorption
        integer(dcm_i4b), intent(in) :: ip, jp, kp, l_h, r_h, nq, nmolecule, nweights, &
            npress, ntemp, nh2o
        integer(dcm_i4b), intent(in), dimension(1-r_h:kp+r_h) :: itemp, ipress
        real(dcm_wp), intent(in), dimension(1:nbands, 1:nmolecule, 1:nweights, 1:ntemp, &
            1:npress, 1:nh2o) :: bli_read
        real(dcm_wp), intent(in), dimension(1:nweights) :: probs_read
        real(dcm_wp), intent(in), dimension(1:nh2o) :: h2o_read
        integer(dcm_i4b), intent(in), dimension(1:nmolecule) :: moleculeID
        real(dcm_wp), intent(in), dimension(1:nmolecule) :: moleculePPM, molecularWeights
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: dz, dzn
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h), &
            intent(inout) :: th, sth
        real(dcm_wp), dimension(-r_h+1:kp+r_h, -r_h+1:jp+r_h, -r_h+1:ip+r_h, nq), &
            intent(in) :: q
        real(dcm_wp), dimension(-l_h+1:kp+r_h), intent(in) :: tref, trefn, rhoa, rhoan
        integer(dcm_i4b), intent(in) :: sub_vert_comm
        real(dcm_wp), intent(in) :: time
        real(dcm_wp), dimension(nbands) :: lambda, b_s_g, lambda_low, lambda_high, &
            delta_lambda, nrwbin, niwbin, sflux_l
        real(dcm_wp), intent(in), dimension(nprocv) :: mvrecv
        integer(dcm_i4b), intent(in) :: nbands, ns, nl, tdstart, tdend
        integer(dcm_i4b), intent(in) :: nprocv
        real(dcm_wp), intent(inout), dimension(1:tdend) :: a, b, c, r, usol
        real(dcm_wp), dimension(1-r_h:kp+r_h, 1-r_h:jp+r_h, 1-r_h:ip+r_h, 1:nbands), &
            intent(inout) :: flux_d, flux_u
        real(dcm_wp), intent(in) :: lat, lon, albedo, emiss, asymmetry_water
        integer(dcm_i4b), intent(in) :: quad_flag, start_year, start_mon, start_day, &
            start_hour, start_min, start_sec, nrad
        real(dcm_wp), dimension(1-r_h:kp+r_h, 1-r_h:jp+r_h, 1-r_h:ip+r_h, 1:nrad), &
            intent(in) :: ngs, lamgs, mugs
        print *, 'dcm: radiative_transfer is not on the GPU path'
        stop
    end subroutine radiative_transfer

end module d_solver


!> module diagnostics (diagnostics.f90:5-17)
module diagnostics
    use iso_c_binding
    use dcm_hip
    use dcm_dropin_util
    implicit none
    private
    public :: horizontal_means, divergence_calc

contains

    !> horizontal_means (diagnostics.f90:31-107): per-level means of the
    !> interior; qbar only when moisture is on
    subroutine horizontal_means(comm3d, id, dims, coords, moisture, ip, jp, ipp, jpp, kpp, &
            l_h, r_h, nq, ubar, vbar, wbar, thbar, qbar, u, v, w, th, q)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        logical, intent(in) :: moisture
        integer(dcm_i4b), intent(in) :: ip, jp, ipp, jpp, kpp, l_h, r_h, nq
        real(dcm_wp), dimension(1-l_h:kpp+r_h), intent(inout) :: ubar, vbar, wbar, thbar
        real(dcm_wp), dimension(1-l_h:kpp+r_h, 1:nq), intent(inout) :: qbar
        real(dcm_wp), dimension(-l_h+1:kpp+r_h, -l_h+1:jpp+r_h, -l_h+1:ipp+r_h), &
            intent(in), target :: u, v, w, th
        real(dcm_wp), dimension(-l_h+1:kpp+r_h, -l_h+1:jpp+r_h, -l_h+1:ipp+r_h, 1:nq), &
            intent(in), target :: q
        type(c_ptr) :: ctx, du, dv, dw, dth, dq
        integer(c_size_t) :: nb
        integer(c_int) :: nqm

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ctx = dcm_context(ipp, jpp, kpp, ones(ipp), ones(jpp), ones(kpp), ones(ipp), &
            ones(jpp), ones(kpp), ones(kpp), ones(kpp), metric_free=.true.)
        nb = field_bytes(ipp, jpp, kpp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        dth = dcm_put(ctx, 13, c_loc(th), nb)
        nqm = 0
        dq = c_null_ptr
        if (moisture .and. nq > 0) then
            nqm = int(nq, c_int)
            dq = dcm_put(ctx, 16, c_loc(q), nb * nq)
            call dcm_check(dcm_hmeans(ctx, du, dv, dw, dth, dq, nqm, ubar, vbar, wbar, thbar, &
                qbar), 'horizontal_means')
        else
            call dcm_check(dcm_hmeans(ctx, du, dv, dw, dth, dq, nqm, ubar, vbar, wbar, thbar, &
                thbar), 'horizontal_means')
        endif
    end subroutine horizontal_means

    !> divergence_calc (diagnostics.f90:120-160): the anelastic divergence of
    !> (u, v, w) on the interior, halos filled like the reference
    subroutine divergence_calc(comm3d, id, dims, coords, ipp, jpp, kpp, dx, dxn, dy, dyn, dz, &
            dzn, rhoa, rhoan, l_h, r_h, u, v, w, div)
        integer(dcm_i4b), intent(in) :: id, comm3d
        integer(dcm_i4b), dimension(3), intent(in) :: dims, coords
        integer(dcm_i4b), intent(in) :: ipp, jpp, kpp, l_h, r_h
        real(dcm_wp), dimension(-l_h+1:ipp+r_h), intent(in) :: dx, dxn
        real(dcm_wp), dimension(-l_h+1:jpp+r_h), intent(in) :: dy, dyn
        real(dcm_wp), dimension(-l_h+1:kpp+r_h), intent(in) :: dz, dzn, rhoa, rhoan
        real(dcm_wp), dimension(-l_h+1:kpp+r_h, -l_h+1:jpp+r_h, -l_h+1:ipp+r_h), &
            intent(in), target :: u, v, w
        real(dcm_wp), dimension(-l_h+1:kpp+r_h, -l_h+1:jpp+r_h, -l_h+1:ipp+r_h), &
            intent(inout), target :: div
        type(c_ptr) :: ctx, du, dv, dw, ddiv
        integer(c_size_t) :: nb

        call need_halo1(l_h, r_h)
        call one_rank(dims)
        ctx = dcm_context(ipp, jpp, kpp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan)
        nb = field_bytes(ipp, jpp, kpp)
        du = dcm_put(ctx, 1, c_loc(u), nb)
        dv = dcm_put(ctx, 2, c_loc(v), nb)
        dw = dcm_put(ctx, 3, c_loc(w), nb)
        ddiv = dcm_put(ctx, 18, c_loc(div), nb)
        call dcm_check(dcm_divergence(ctx, du, dv, dw, ddiv), 'divergence_calc')
        call dcm_take(ctx, 18, c_loc(div), nb)
    end subroutine divergence_calc

end module diagnostics
