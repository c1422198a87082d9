!> Calls the drop-in modules in model_driver order (driver_code.f90:433-580)
!> on one input set and writes every result, so that a test can compare the
!> Fortran boundary end to end with a CPU restatement of the same sequence.
!>
!> usage: dcm_dropin_check <input file> <output file>
!> input (stream, native byte order): int32 ip, jp, kp, nq, kord, monotone;
!> float64 dt, tol; 1-D dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, theta,
!> thetan, xg, yg, zg (each n+2); padded fields u, v, w, th, q(nq), p;
!> float64 z0, z0th, forcing_tau; lamsq, dampfacn, dampfac, u_force, v_force,
!> w_subs (each kp+2).
!> output: thbar, qbar(nq), div, su, sv, sw, sth, rhs, p, u, v, w (advanced),
!> th (after adv_ref_state and mpdata_3d), q (after mpdata_vec_3d),
!> psi (first_order_upstream_3d of the input th), then on the input fields:
!> thv (mpdata_vert_3d of th with w), qsed (mpdata_vec_vert_3d of q with w,
!> as PAMM calls it: kord 1, no limiter, neumann 2), tha (mpdata_3d_add of th
!> with the base-state profile theta), then sources with every option but
!> radiation on (the advanced u, v, w as the current level, the input ones
!> as the previous, the input th and q, the means of the input fields):
!> su, sv, sw, sth, rhs, sq(nq), strain, vism, vist, viss(nq).
program dropin_check
    use iso_c_binding
    use advection_s_3d, only: mpdata_3d, mpdata_vec_3d, first_order_upstream_3d, adv_ref_state, &
        mpdata_vert_3d, mpdata_vec_vert_3d, mpdata_3d_add
    use d_solver, only: bicgstab, sources, advance_momentum
    use diagnostics, only: horizontal_means, divergence_calc
    implicit none
    integer, parameter :: wp = c_double, i4b = c_int32_t
    integer(i4b) :: ip, jp, kp, nq, kord, imono, l_h, r_h
    integer(i4b) :: dims(3), coords(3), comm3d, id, rank
    real(wp) :: dt, tol, lbc, ubc, z0, z0th, forcing_tau
    real(wp), allocatable, dimension(:) :: dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, theta, &
        thetan, xg, yg, zg, ubar, vbar, wbar, thbar, zero_k, lbcq, ubcq, lamsq, dampfacn, &
        dampfac, u_force, v_force, w_subs
    real(wp), allocatable :: qbar(:, :)
    real(wp), allocatable, dimension(:, :, :) :: u, v, w, th, p, zu, zv, zw, su, sv, sw, sth, &
        rhs, strain, vism, vist, div, psi, rad_power, thv, tha, u0, v0, w0, th0, su2, sv2, &
        sw2, sth2, rhs2, strain2, vism2, vist2
    real(wp), allocatable, dimension(:, :, :, :) :: q, sq, viss, flux_u, flux_d, ngs, lamgs, &
        mugs, qsed, q0, sq2, viss2
    real(wp) :: rad1(1), mvrecv(1), tri(1)
    character(len=1024) :: fin, fout
    integer :: iu
    logical :: monotone

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old')
    read(iu) ip, jp, kp, nq, kord, imono
    monotone = imono /= 0
    l_h = 1
    r_h = 1
    dims = 1
    coords = 0
    comm3d = 0
    id = 0
    rank = 0
    read(iu) dt, tol
    allocate(dx(0:ip+1), dy(0:jp+1), dz(0:kp+1), dxn(0:ip+1), dyn(0:jp+1), dzn(0:kp+1), &
        rhoa(0:kp+1), rhoan(0:kp+1), theta(0:kp+1), thetan(0:kp+1), xg(0:ip+1), yg(0:jp+1), &
        zg(0:kp+1))
    read(iu) dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, theta, thetan, xg, yg, zg
    ! (a MOLD must already be allocated: give u its shape first)
    allocate(u(0:kp+1, 0:jp+1, 0:ip+1))
    allocate(v, w, th, p, zu, zv, zw, su, sv, sw, sth, rhs, strain, vism, vist, div, psi, &
        rad_power, thv, tha, u0, v0, w0, th0, su2, sv2, sw2, sth2, rhs2, strain2, vism2, &
        vist2, mold=u)
    allocate(q(0:kp+1, 0:jp+1, 0:ip+1, nq))
    allocate(sq, viss, qsed, q0, sq2, viss2, mold=q)
    read(iu) u, v, w, th, q, p
    allocate(lamsq(0:kp+1), dampfacn(0:kp+1), dampfac(0:kp+1), u_force(0:kp+1), &
        v_force(0:kp+1), w_subs(0:kp+1))
    read(iu) z0, z0th, forcing_tau
    read(iu) lamsq, dampfacn, dampfac, u_force, v_force, w_subs
    close(iu)
    u0 = u
    v0 = v
    w0 = w
    th0 = th
    q0 = q
    thv = th
    tha = th
    qsed = q
    allocate(ubar(0:kp+1), vbar(0:kp+1), wbar(0:kp+1), thbar(0:kp+1), zero_k(0:kp+1), &
        qbar(0:kp+1, nq), lbcq(nq), ubcq(nq))
    allocate(flux_u(0:kp+1, 0:jp+1, 0:ip+1, 1))
    allocate(flux_d, ngs, lamgs, mugs, mold=flux_u)
    zero_k = 0.0_wp
    zu = u
    zv = v
    zw = w
    su = 0.0_wp
    sv = 0.0_wp
    sw = 0.0_wp
    rhs = 0.0_wp
    sth = 0.0_wp
    sq = 0.0_wp
    strain = 0.0_wp
    vism = 0.0_wp
    vist = 0.0_wp
    viss = 0.0_wp
    rad_power = 0.0_wp
    flux_u = 0.0_wp
    flux_d = 0.0_wp
    ngs = 0.0_wp
    lamgs = 0.0_wp
    mugs = 0.0_wp
    div = 0.0_wp
    rad1 = 0.0_wp
    mvrecv = 0.0_wp
    tri = 0.0_wp

    call horizontal_means(comm3d, id, dims, coords, .true., ip, jp, ip, jp, kp, l_h, r_h, nq, &
        ubar, vbar, wbar, thbar, qbar, u, v, w, th, q)
    call divergence_calc(comm3d, id, dims, coords, ip, jp, kp, dx, dxn, dy, dyn, dz, dzn, &
        rhoa, rhoan, l_h, r_h, u, v, w, div)
    call sources(comm3d, id, rank, dims, coords, dt, xg, yg, zg, zg, dx, dy, dz, dxn, dyn, &
        dzn, ip, jp, kp, l_h, r_h, nq, ubar, vbar, wbar, thbar, qbar, zero_k, zero_k, &
        zero_k, zero_k, 0.0_wp, zero_k, zu, zv, zw, u, v, w, su, sv, sw, q, sq, viss, rhs, &
        th, sth, strain, vism, vist, theta, thetan, theta, thetan, rhoa, rhoan, zero_k, &
        zero_k, 0.0_wp, 0.0_wp, .false., .true., .false., .false., .false., .false., 0, &
        0.0_wp, 1, 1, tri, tri, tri, tri, tri, 1, 1, 1, flux_u, flux_d, rad_power, rad1, &
        rad1, rad1, rad1, rad1, rad1, rad1, rad1, 2000, 1, 1, 0, 0, 0, 0.0_wp, 0.0_wp, &
        0.0_wp, 0.0_wp, 0, 0.0_wp, 1, ngs, lamgs, mugs, 1, mvrecv)
    call report('sources su', su)
    call report('sources sv', sv)
    call report('sources sw', sw)
    call report('sources rhs', rhs(1:kp, 1:jp, 1:ip))
    call report('p (input)', p)
    call lateral_faces(rhs)
    call bicgstab(comm3d, id, rank, dims, coords, dt, xg, yg, zg, dx, dy, dz, dxn, dyn, dzn, &
        rhoa, rhoan, ip, jp, kp, l_h, r_h, su, sv, sw, zu, zv, zw, p, rhs, tol, .false.)
    call lateral_faces(p)
    ! the first order scheme and the reference state on the input fields
    psi = th
    call first_order_upstream_3d(dt, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, u, v, &
        w, psi, 0, dims, coords)
    call adv_ref_state(dt, dx, dx, dz, dxn, dxn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, u, v, &
        w, th, thetan, comm3d, id, dims, coords)
    lbc = 0.0_wp
    ubc = 0.0_wp
    call mpdata_3d(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, u, v, w, &
        th, lbc, ubc, kord, monotone, 0, comm3d, id, dims, coords)
    lbcq = 0.0_wp
    ubcq = 0.0_wp
    call mpdata_vec_3d(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, nq, l_h, r_h, &
        u, v, w, q, lbcq, ubcq, kord, monotone, 0, comm3d, id, dims, coords)
    call advance_momentum(comm3d, id, rank, 2.0_wp * dt, dx, dy, dz, dxn, dyn, dzn, rhoa, &
        rhoan, ip, jp, kp, l_h, r_h, u, v, w, zu, zv, zw, su, sv, sw, p, dims, coords)
    ! vertical-only and profile-added advection of the input fields
    lbc = 0.0_wp
    ubc = 0.0_wp
    call mpdata_vert_3d(dt, dz, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, zw, thv, lbc, ubc, &
        kord, monotone, 0, comm3d, id, dims, coords)
    lbcq = 0.0_wp
    ubcq = 0.0_wp
    call mpdata_vec_vert_3d(dt, dz, dzn, rhoa, rhoan, ip, jp, kp, nq, l_h, r_h, zw, qsed, lbcq, &
        ubcq, 1, .false., 2, comm3d, id, dims, coords)
    lbc = 0.0_wp
    ubc = 0.0_wp
    call mpdata_3d_add(dt, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, ip, jp, kp, l_h, r_h, zu, zv, &
        zw, tha, theta, lbc, ubc, kord, monotone, 0, comm3d, id, dims, coords)

    ! every option of sources but radiation (the means are of the input fields)
    su2 = 0.0_wp
    sv2 = 0.0_wp
    sw2 = 0.0_wp
    sth2 = 0.0_wp
    rhs2 = 0.0_wp
    sq2 = 0.0_wp
    strain2 = 0.0_wp
    vism2 = 0.0_wp
    vist2 = 0.0_wp
    viss2 = 0.0_wp
    call sources(comm3d, id, rank, dims, coords, dt, xg, yg, zg, zg, dx, dy, dz, dxn, dyn, &
        dzn, ip, jp, kp, l_h, r_h, nq, ubar, vbar, wbar, thbar, qbar, dampfacn, dampfac, &
        u_force, v_force, forcing_tau, w_subs, u0, v0, w0, u, v, w, su2, sv2, sw2, q0, sq2, &
        viss2, rhs2, th0, sth2, strain2, vism2, vist2, theta, thetan, theta, thetan, rhoa, &
        rhoan, lamsq, lamsq, z0, z0th, .true., .true., .true., .true., .true., .false., 0, &
        0.0_wp, 1, 1, tri, tri, tri, tri, tri, 1, 1, 1, flux_u, flux_d, rad_power, rad1, &
        rad1, rad1, rad1, rad1, rad1, rad1, rad1, 2000, 1, 1, 0, 0, 0, 0.0_wp, 0.0_wp, &
        0.0_wp, 0.0_wp, 0, 0.0_wp, 1, ngs, lamgs, mugs, 1, mvrecv)

    open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(iu) thbar, qbar, div, su, sv, sw, sth, rhs, p, u, v, w, th, q, psi, thv, qsed, tha
    write(iu) su2, sv2, sw2, sth2, rhs2, sq2, strain2, vism2, vist2, viss2
    close(iu)

contains

    !> print where an array holds non-finite values (nothing when it is clean)
    subroutine report(what, a)
        use, intrinsic :: ieee_arithmetic, only: ieee_is_finite
        character(len=*), intent(in) :: what
        real(wp), intent(in) :: a(:, :, :)
        integer :: nbad
        nbad = count(.not. ieee_is_finite(a))
        if (nbad > 0) print *, what, ': ', nbad, ' non-finite of ', size(a), &
            ' first at ', findloc(.not. ieee_is_finite(a), .true.)
    end subroutine report

    !> exchange_along_dim_wo on one rank (mpi_module.f90:1285-1494): the
    !> x and y faces, interior rows only
    subroutine lateral_faces(a)
        real(wp), intent(inout) :: a(0:kp+1, 0:jp+1, 0:ip+1)
        a(1:kp, 1:jp, 0) = a(1:kp, 1:jp, ip)
        a(1:kp, 1:jp, ip+1) = a(1:kp, 1:jp, 1)
        a(1:kp, 0, 1:ip) = a(1:kp, jp, 1:ip)
        a(1:kp, jp+1, 1:ip) = a(1:kp, 1, 1:ip)
    end subroutine lateral_faces

end program dropin_check
