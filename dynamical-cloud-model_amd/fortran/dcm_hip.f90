!> ISO_C_BINDING interface of libdcm_hip.so (include/dcm_hip.h) and the
!> helpers the Fortran drop-in modules share: one device context per grid,
!> device buffers per argument slot, and error reporting in the reference's print/stop style.
module dcm_hip
    use iso_c_binding
    implicit none
    private

    !> kinds of the reference's FP64 build (numerics_type.f90, VAR_TYPE=1)
    integer, parameter, public :: dcm_wp = c_double, dcm_i4b = c_int32_t

    integer(c_int), parameter, public :: DCM_OK = 0, DCM_NOT_CONVERGED = 1, &
        DCM_BREAKDOWN = 2, DCM_NONFINITE = 3
    integer(c_int), parameter, public :: DCM_FP64 = 0, DCM_MIXED = 1
    integer(c_int), parameter, public :: DCM_SUM_TREE = 0, DCM_SUM_SEQUENTIAL = 1
    integer(c_int), parameter, public :: DCM_F_U = 0, DCM_F_V = 1, DCM_F_W = 2, &
        DCM_F_ZU = 3, DCM_F_ZV = 4, DCM_F_ZW = 5, DCM_F_TH = 6, DCM_F_P = 7, &
        DCM_F_PSRC = 8, DCM_F_SU = 9, DCM_F_SV = 10, DCM_F_SW = 11, DCM_F_STH = 12, &
        DCM_F_Q = 13, DCM_F_THBAR = 14

    !> options of sources beyond the inviscid branch (C dcm_physics); the
    !> pointers are c_loc of host arrays of kp+2, or c_null_ptr
    type, bind(c), public :: dcm_physics
        integer(c_int) :: viscous = 0, damping_layer = 0, forcing = 0, divergence = 0
        real(c_double) :: forcing_tau = 1.0_c_double, z0 = 1.0e-4_c_double, &
            z0th = 1.0e-4_c_double
        type(c_ptr) :: lamsq = c_null_ptr, dampfacn = c_null_ptr, dampfac = c_null_ptr, &
            u_force = c_null_ptr, v_force = c_null_ptr, w_subs = c_null_ptr, &
            z = c_null_ptr, theta = c_null_ptr, thetan = c_null_ptr
    end type dcm_physics

    type, bind(c), public :: dcm_step_stats
        integer(c_int) :: steps, iters_total, iters_max, outer_total, fallbacks, reserved
        real(c_double) :: last_err, max_true_resid
    end type dcm_step_stats

    public :: dcm_context, dcm_check, dcm_put, dcm_take, &
        dcm_slot, field_bytes, need_halo1
    public :: dcm_malloc, dcm_free, dcm_memcpy_htod, dcm_memcpy_dtoh
    public :: dcm_context_create, dcm_context_destroy, dcm_synchronize, dcm_mirror, &
        dcm_mirror_upload, dcm_mirror_download, dcm_mirror_release_all, dcm_last_error, &
        dcm_fill_full, dcm_fill_along_dim, dcm_fill_along_dim_wo, dcm_fill_full_wo, &
        dcm_fill_along_z, dcm_upwind3d, dcm_mpdata3d, dcm_mpdata_vec3d, dcm_adv_ref_state, &
        dcm_mpdata3d_add, dcm_upwind_vert3d, dcm_mpdata_vert3d, dcm_mpdata_vec_vert3d, &
        dcm_subgrid, dcm_sources_phys, dcm_model_set_physics, &
        dcm_hmeans, dcm_sources, dcm_divergence, dcm_set_sum_order, dcm_set_precond_chebyshev, &
        dcm_set_mpdata_variant, dcm_set_precond_multigrid, &
        dcm_bicgstab, &
        dcm_advance_momentum, dcm_advance_scalars, &
        dcm_model_create, dcm_model_destroy, dcm_model_set_precision, dcm_model_set, &
        dcm_model_get, dcm_model_init, dcm_model_step, &
        dcm_ensemble_create, dcm_ensemble_destroy, dcm_ensemble_autotune, dcm_ensemble_step, &
        dcm_ensemble_set_graphs, &
        dcm_nc_create, dcm_nc_put_record, dcm_nc_close, dcm_output_create, dcm_output_write, &
        dcm_output_flush, dcm_output_stats, dcm_output_destroy

    interface
        integer(c_int) function dcm_context_create(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, &
                rhoa, rhoan, ctx) bind(c, name="dcm_context_create")
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: ip, jp, kp
            real(c_double), intent(in) :: dx(*), dy(*), dz(*), dxn(*), dyn(*), dzn(*), &
                rhoa(*), rhoan(*)
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function dcm_context_destroy(ctx) bind(c, name="dcm_context_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dcm_synchronize(ctx) bind(c, name="dcm_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dcm_malloc(ctx, bytes, dptr) bind(c, name="dcm_malloc")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx
            integer(c_size_t), value :: bytes
            type(c_ptr), intent(out) :: dptr
        end function
        integer(c_int) function dcm_free(ctx, dptr) bind(c, name="dcm_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, dptr
        end function
        integer(c_int) function dcm_memcpy_htod(ctx, dst, src, bytes) bind(c, name="dcm_memcpy_htod")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, dst, src
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function dcm_memcpy_dtoh(ctx, dst, src, bytes) bind(c, name="dcm_memcpy_dtoh")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, dst, src
            integer(c_size_t), value :: bytes
        end function
        type(c_ptr) function dcm_last_error() bind(c, name="dcm_last_error")
            import :: c_ptr
        end function
        integer(c_int) function dcm_mirror(ctx, host, bytes, dptr) bind(c, name="dcm_mirror")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, host
            integer(c_size_t), value :: bytes
            type(c_ptr), intent(out) :: dptr
        end function
        integer(c_int) function dcm_mirror_upload(ctx, host) bind(c, name="dcm_mirror_upload")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, host
        end function
        integer(c_int) function dcm_mirror_download(ctx, host) bind(c, name="dcm_mirror_download")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, host
        end function
        integer(c_int) function dcm_mirror_release_all(ctx) bind(c, name="dcm_mirror_release_all")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function dcm_fill_full(ctx, a, lbc, ubc) bind(c, name="dcm_fill_full")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, a
            real(c_double), value :: lbc, ubc
        end function
        integer(c_int) function dcm_fill_along_dim(ctx, a, lbc, ubc) bind(c, name="dcm_fill_along_dim")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, a
            real(c_double), value :: lbc, ubc
        end function
        integer(c_int) function dcm_fill_along_dim_wo(ctx, a) bind(c, name="dcm_fill_along_dim_wo")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, a
        end function
        integer(c_int) function dcm_fill_full_wo(ctx, a) bind(c, name="dcm_fill_full_wo")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, a
        end function
        integer(c_int) function dcm_fill_along_z(ctx, a, lbc, ubc) bind(c, name="dcm_fill_along_z")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, a
            real(c_double), value :: lbc, ubc
        end function
        integer(c_int) function dcm_upwind3d(ctx, dt, u, v, w, psi, neumann) &
                bind(c, name="dcm_upwind3d")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, psi
            real(c_double), value :: dt
            integer(c_int), value :: neumann
        end function
        integer(c_int) function dcm_mpdata3d(ctx, dt, u, v, w, psi, lbc, ubc, kord, monotone, &
                neumann, early_exit) bind(c, name="dcm_mpdata3d")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, psi
            real(c_double), value :: dt
            real(c_double), intent(inout) :: lbc, ubc
            integer(c_int), value :: kord, monotone, neumann
            integer(c_int), intent(out) :: early_exit
        end function
        integer(c_int) function dcm_mpdata_vec3d(ctx, dt, u, v, w, psi, nq, lbc, ubc, kord, &
                monotone, neumann, early_exit) bind(c, name="dcm_mpdata_vec3d")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, psi
            real(c_double), value :: dt
            integer(c_int), value :: nq, kord, monotone, neumann
            real(c_double), intent(inout) :: lbc(*), ubc(*)
            integer(c_int), intent(out) :: early_exit
        end function
        integer(c_int) function dcm_subgrid(ctx, ph, u, v, w, zu, zv, zw, th, q, nq, vism, vist, &
                viss, strain, su, sv, sw, sth, sq) bind(c, name="dcm_subgrid")
            import :: c_int, c_ptr, dcm_physics
            type(c_ptr), value :: ctx, u, v, w, zu, zv, zw, th, q
            type(dcm_physics), intent(in) :: ph
            integer(c_int), value :: nq
            type(c_ptr), value :: vism, vist, viss, strain, su, sv, sw, sth, sq
        end function
        integer(c_int) function dcm_sources_phys(ctx, ph, ubar, vbar, wbar, u, v, w, zu, zv, zw, &
                th, q, nq, theta, thbar, su, sv, sw, sth, sq, strain, vism, vist, viss, rhs) &
                bind(c, name="dcm_sources_phys")
            import :: c_int, c_ptr, c_double, dcm_physics
            type(c_ptr), value :: ctx
            type(dcm_physics), intent(in) :: ph
            real(c_double), intent(in) :: ubar(*), vbar(*), wbar(*), theta(*), thbar(*)
            type(c_ptr), value :: u, v, w, zu, zv, zw, th, q
            integer(c_int), value :: nq
            type(c_ptr), value :: su, sv, sw, sth, sq, strain, vism, vist, viss, rhs
        end function
        integer(c_int) function dcm_mpdata3d_add(ctx, dt, u, v, w, psi, psi_1d, lbc, ubc, kord, &
                monotone, neumann, early_exit) bind(c, name="dcm_mpdata3d_add")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, psi
            real(c_double), value :: dt
            real(c_double), intent(in) :: psi_1d(*)
            real(c_double), intent(inout) :: lbc, ubc
            integer(c_int), value :: kord, monotone, neumann
            integer(c_int), intent(out) :: early_exit
        end function
        integer(c_int) function dcm_upwind_vert3d(ctx, dt, w, psi, neumann) &
                bind(c, name="dcm_upwind_vert3d")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, w, psi
            real(c_double), value :: dt
            integer(c_int), value :: neumann
        end function
        integer(c_int) function dcm_mpdata_vert3d(ctx, dt, w, psi, lbc, ubc, kord, monotone, &
                neumann, early_exit) bind(c, name="dcm_mpdata_vert3d")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, w, psi
            real(c_double), value :: dt
            real(c_double), intent(inout) :: lbc, ubc
            integer(c_int), value :: kord, monotone, neumann
            integer(c_int), intent(out) :: early_exit
        end function
        integer(c_int) function dcm_mpdata_vec_vert3d(ctx, dt, w, psi, nq, lbc, ubc, kord, &
                monotone, neumann, early_exit) bind(c, name="dcm_mpdata_vec_vert3d")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, w, psi
            real(c_double), value :: dt
            integer(c_int), value :: nq, kord, monotone, neumann
            real(c_double), intent(inout) :: lbc(*), ubc(*)
            integer(c_int), intent(out) :: early_exit
        end function
        integer(c_int) function dcm_adv_ref_state(ctx, dt, w, psi, psi_ref) &
                bind(c, name="dcm_adv_ref_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, w, psi
            real(c_double), value :: dt
            real(c_double), intent(in) :: psi_ref(*)
        end function
        integer(c_int) function dcm_hmeans(ctx, u, v, w, th, q, nq, ubar, vbar, wbar, thbar, &
                qbar) bind(c, name="dcm_hmeans")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, th, q
            integer(c_int), value :: nq
            real(c_double), intent(out) :: ubar(*), vbar(*), wbar(*), thbar(*), qbar(*)
        end function
        integer(c_int) function dcm_sources(ctx, u, v, w, th, theta, thbar, su, sv, sw, sth, &
                rhs) bind(c, name="dcm_sources")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, th, su, sv, sw, sth, rhs
            real(c_double), intent(in) :: theta(*), thbar(*)
        end function
        integer(c_int) function dcm_set_sum_order(ctx, order) bind(c, name="dcm_set_sum_order")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: order
        end function
        ! the mixed inner solve's preconditioner: one multigrid V-cycle per
        ! application (on /= 0) instead of the Jacobi sweeps
        integer(c_int) function dcm_set_precond_multigrid(ctx, on) &
                bind(c, name="dcm_set_precond_multigrid")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function
        ! the copy of advection_3d.F90 MPDATA follows: 0 pamm/sfvt (linked
        ! into main.exe, the default), 1 sfvt
        integer(c_int) function dcm_set_mpdata_variant(ctx, variant) &
                bind(c, name="dcm_set_mpdata_variant")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: variant
        end function
        ! the mixed inner solve's preconditioner: Chebyshev factors on [a, 2]
        ! with `divisions` applications of 1/a_p (a = 0: the reference's)
        integer(c_int) function dcm_set_precond_chebyshev(ctx, a, divisions) &
                bind(c, name="dcm_set_precond_chebyshev")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: a
            integer(c_int), value :: divisions
        end function
        integer(c_int) function dcm_divergence(ctx, u, v, w, div) bind(c, name="dcm_divergence")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, div
        end function
        integer(c_int) function dcm_bicgstab(ctx, xg, yg, zg, sw, x, b, tol, test_solver, &
                itmax, iters, err) bind(c, name="dcm_bicgstab")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, sw, x, b
            real(c_double), intent(in) :: xg(*), yg(*), zg(*)
            real(c_double), value :: tol
            integer(c_int), value :: test_solver, itmax
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: err
        end function
        integer(c_int) function dcm_advance_momentum(ctx, dt, u, v, w, zu, zv, zw, su, sv, sw, &
                p) bind(c, name="dcm_advance_momentum")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, zu, zv, zw, su, sv, sw, p
            real(c_double), value :: dt
        end function
        integer(c_int) function dcm_advance_scalars(ctx, dt, q, sq, th, sth, nq, moisture) &
                bind(c, name="dcm_advance_scalars")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, q, sq, th, sth
            real(c_double), value :: dt
            integer(c_int), value :: nq, moisture
        end function
        integer(c_int) function dcm_model_create(ctx, nq, ncat, c_s, c_e, dt, kord, monotone, &
                advection_scheme, ptol, theta, thetan, xg, yg, zg, model) &
                bind(c, name="dcm_model_create")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nq, ncat, kord, monotone, advection_scheme
            integer(c_int), intent(in) :: c_s(*), c_e(*)
            real(c_double), value :: dt, ptol
            real(c_double), intent(in) :: theta(*), thetan(*), xg(*), yg(*), zg(*)
            type(c_ptr), intent(out) :: model
        end function
        integer(c_int) function dcm_model_destroy(model) bind(c, name="dcm_model_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: model
        end function
        integer(c_int) function dcm_model_set_precision(model, precision) &
                bind(c, name="dcm_model_set_precision")
            import :: c_int, c_ptr
            type(c_ptr), value :: model
            integer(c_int), value :: precision
        end function
        integer(c_int) function dcm_model_set(model, field, idx, host) bind(c, name="dcm_model_set")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: model
            integer(c_int), value :: field, idx
            real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function dcm_model_get(model, field, idx, host) bind(c, name="dcm_model_get")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: model
            integer(c_int), value :: field, idx
            real(c_double), intent(out) :: host(*)
        end function
        integer(c_int) function dcm_model_init(model) bind(c, name="dcm_model_init")
            import :: c_int, c_ptr
            type(c_ptr), value :: model
        end function
        integer(c_int) function dcm_model_set_physics(model, ph) bind(c, name="dcm_model_set_physics")
            import :: c_int, c_ptr, dcm_physics
            type(c_ptr), value :: model
            type(dcm_physics), intent(in) :: ph
        end function
        integer(c_int) function dcm_model_step(model, nsteps, stats) bind(c, name="dcm_model_step")
            import :: c_int, c_ptr, dcm_step_stats
            type(c_ptr), value :: model
            integer(c_int), value :: nsteps
            type(dcm_step_stats), intent(out) :: stats
        end function
        ! ensembles of independent models (each with its own context)
        integer(c_int) function dcm_ensemble_create(n, models, ens) &
                bind(c, name="dcm_ensemble_create")
            import :: c_int, c_ptr
            integer(c_int), value :: n
            type(c_ptr), intent(in) :: models(*)
            type(c_ptr), intent(out) :: ens
        end function
        integer(c_int) function dcm_ensemble_destroy(ens) bind(c, name="dcm_ensemble_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ens
        end function
        integer(c_int) function dcm_ensemble_autotune(ens, march_planes) &
                bind(c, name="dcm_ensemble_autotune")
            import :: c_int, c_ptr
            type(c_ptr), value :: ens
            integer(c_int), intent(out) :: march_planes
        end function
        integer(c_int) function dcm_ensemble_step(ens, nsteps, stats) &
                bind(c, name="dcm_ensemble_step")
            import :: c_int, c_ptr, dcm_step_stats
            type(c_ptr), value :: ens
            integer(c_int), value :: nsteps
            type(dcm_step_stats), intent(out) :: stats(*)
        end function
        integer(c_int) function dcm_ensemble_set_graphs(ens, enable) &
                bind(c, name="dcm_ensemble_set_graphs")
            import :: c_int, c_ptr
            type(c_ptr), value :: ens
            integer(c_int), value :: enable
        end function
        ! the output file (driver_code.f90:725-1145); path is null-terminated
        ! (trim(outputfile)//c_null_char), q_names the character(len=20)
        ! array, precip c_loc of the (1:kp,0:jp+1,0:ip+1,nprec) array or
        ! c_null_ptr
        integer(c_int) function dcm_nc_create(path, ip, jp, kp, moisture, nq, nprec, q_names, &
                x, y, z, rhoa, thetan, trefn, f) bind(c, name="dcm_nc_create")
            import :: c_int, c_double, c_ptr, c_char
            character(kind=c_char), intent(in) :: path(*), q_names(*)
            integer(c_int), value :: ip, jp, kp, moisture, nq, nprec
            real(c_double), intent(in) :: x(*), y(*), z(*), rhoa(*), thetan(*), trefn(*)
            type(c_ptr), intent(out) :: f
        end function
        integer(c_int) function dcm_nc_put_record(f, n, time, u, v, w, th, p, div, q, precip) &
                bind(c, name="dcm_nc_put_record")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: f
            integer(c_int), value :: n
            real(c_double), value :: time
            real(c_double), intent(in) :: u(*), v(*), w(*), th(*), p(*), div(*), q(*)
            type(c_ptr), value :: precip
        end function
        integer(c_int) function dcm_nc_close(f) bind(c, name="dcm_nc_close")
            import :: c_int, c_ptr
            type(c_ptr), value :: f
        end function
        integer(c_int) function dcm_output_create(model, path, moisture, nprec, q_names, x, y, &
                z, rhoa, thetan, trefn, o) bind(c, name="dcm_output_create")
            import :: c_int, c_double, c_ptr, c_char
            type(c_ptr), value :: model
            character(kind=c_char), intent(in) :: path(*), q_names(*)
            integer(c_int), value :: moisture, nprec
            real(c_double), intent(in) :: x(*), y(*), z(*), rhoa(*), thetan(*), trefn(*)
            type(c_ptr), intent(out) :: o
        end function
        integer(c_int) function dcm_output_write(o, time, precip) bind(c, name="dcm_output_write")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: o
            real(c_double), value :: time
            type(c_ptr), value :: precip
        end function
        integer(c_int) function dcm_output_flush(o) bind(c, name="dcm_output_flush")
            import :: c_int, c_ptr
            type(c_ptr), value :: o
        end function
        integer(c_int) function dcm_output_stats(o, records, stall_s, write_s) &
                bind(c, name="dcm_output_stats")
            import :: c_int, c_double, c_ptr, c_long_long
            type(c_ptr), value :: o
            integer(c_long_long), intent(out) :: records
            real(c_double), intent(out) :: stall_s, write_s
        end function
        integer(c_int) function dcm_output_destroy(o) bind(c, name="dcm_output_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: o
        end function
    end interface

    ! one device context per grid shape, created on first use
    type(c_ptr), save :: g_ctx = c_null_ptr
    integer, save :: g_ip = -1, g_jp = -1, g_kp = -1
    logical, save :: g_provisional = .false.
    ! device buffers of the drop-in routines, one per argument slot
    integer, parameter :: nslots = 40
    type(c_ptr), save :: g_slot(nslots) = c_null_ptr
    integer(c_size_t), save :: g_slot_bytes(nslots) = 0

contains

    !> The context for a grid, created on first use; a different grid shape
    !> replaces it.  Metric arrays are the reference's 1-D arrays (0:n+1).
    !> A routine that receives only some metrics (first_order_upstream_3d,
    !> adv_ref_state) passes stand-ins for the others with provisional; one
    !> that reads none (horizontal_means) passes metric_free and reuses any
    !> context of its shape.  A provisional context is replaced by the next
    !> call that brings metrics.  Like the reference's bicgstab, which builds
    !> its coefficients on its first call (dynamics.f90:330-367), the metrics
    !> of a grid are taken once from the first complete set.
    function dcm_context(ip, jp, kp, dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, provisional, &
            metric_free) result(ctx)
        integer, intent(in) :: ip, jp, kp
        real(c_double), intent(in) :: dx(0:ip+1), dy(0:jp+1), dz(0:kp+1), dxn(0:ip+1), &
            dyn(0:jp+1), dzn(0:kp+1), rhoa(0:kp+1), rhoan(0:kp+1)
        logical, intent(in), optional :: provisional, metric_free
        type(c_ptr) :: ctx
        logical :: prov, free
        prov = .false.
        free = .false.
        if (present(provisional)) prov = provisional
        if (present(metric_free)) free = metric_free
        if (c_associated(g_ctx)) then
            if (ip /= g_ip .or. jp /= g_jp .or. kp /= g_kp .or. &
                    (g_provisional .and. .not. free)) then
                call dcm_check(dcm_context_destroy(g_ctx), 'dcm_context_destroy')
                g_ctx = c_null_ptr
            endif
        endif
        if (.not. c_associated(g_ctx)) then
            call dcm_check(dcm_context_create(int(ip, c_int), int(jp, c_int), int(kp, c_int), &
                dx, dy, dz, dxn, dyn, dzn, rhoa, rhoan, g_ctx), 'dcm_context_create')
            g_ip = ip
            g_jp = jp
            g_kp = kp
            g_provisional = prov .or. free
        endif
        ctx = g_ctx
    end function dcm_context

    !> bytes of one padded (0:kp+1,0:jp+1,0:ip+1) FP64 field
    integer(c_size_t) function field_bytes(ip, jp, kp)
        integer, intent(in) :: ip, jp, kp
        field_bytes = int(kp + 2, c_size_t) * int(jp + 2, c_size_t) * int(ip + 2, c_size_t) * 8_c_size_t
    end function field_bytes

    !> the GPU path keeps the reference's halo of one cell (main.f90:80-81)
    subroutine need_halo1(l_h, r_h)
        integer, intent(in) :: l_h, r_h
        if (l_h /= 1 .or. r_h /= 1) then
            print *, 'dcm: the GPU path supports a halo of one cell only'
            error stop 1
        endif
    end subroutine need_halo1

    !> device buffer of an argument slot, at least nbytes
    function dcm_slot(ctx, slot, nbytes) result(d)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: slot
        integer(c_size_t), intent(in) :: nbytes
        type(c_ptr) :: d
        if (g_slot_bytes(slot) < nbytes) then
            if (c_associated(g_slot(slot))) call dcm_check(dcm_free(ctx, g_slot(slot)), 'dcm_free')
            call dcm_check(dcm_malloc(ctx, nbytes, g_slot(slot)), 'dcm_malloc')
            g_slot_bytes(slot) = nbytes
        endif
        d = g_slot(slot)
    end function dcm_slot

    !> upload a host array into a slot buffer
    function dcm_put(ctx, slot, host, nbytes) result(d)
        type(c_ptr), intent(in) :: ctx, host
        integer, intent(in) :: slot
        integer(c_size_t), intent(in) :: nbytes
        type(c_ptr) :: d
        d = dcm_slot(ctx, slot, nbytes)
        call dcm_check(dcm_memcpy_htod(ctx, d, host, nbytes), 'dcm_memcpy_htod')
    end function dcm_put

    !> copy a slot buffer back into a host array
    subroutine dcm_take(ctx, slot, host, nbytes)
        type(c_ptr), intent(in) :: ctx, host
        integer, intent(in) :: slot
        integer(c_size_t), intent(in) :: nbytes
        call dcm_check(dcm_memcpy_dtoh(ctx, host, g_slot(slot), nbytes), 'dcm_memcpy_dtoh')
    end subroutine dcm_take

    !> the reference's failure style: print the message and stop
    subroutine dcm_check(st, what)
        integer(c_int), intent(in) :: st
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (st == DCM_OK) return
        call c_f_pointer(dcm_last_error(), msg, [1024])
        n = 0
        do while (n < 1024)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        enddo
        print *, what, ' failed with status ', st, ' ', msg(1:n)
        error stop 1
    end subroutine dcm_check

end module dcm_hip
