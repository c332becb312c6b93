!  nsx_iface.F90 -- ISO_C_BINDING interfaces of the nsx C ABI (include/nsx.h)
!
!  One interface per entry point of include/nsx.h, with the same argument
!  meaning: arrays are passed by address in the reference layout ((N_CELL,NZ),
!  (N_VERT,NZ-1), (N_CELL0,NZ) column-major, 1-based connectivity), scalars by
!  value, the context as an opaque c_ptr.  Return codes: 0 ok, > 0 a soft
!  solver outcome (NSX_SOLVE_*), < 0 an error whose text nsx_last_error gives.
      module nsx_iface
      use iso_c_binding
      implicit none

!     field ids (enum nsx_field)
      integer(c_int), parameter :: NSX_F_UFN = 0, NSX_F_VFN = 1,        &
           NSX_F_WFN = 2, NSX_F_PFN = 3, NSX_F_UF = 4, NSX_F_VF = 5,     &
           NSX_F_WF = 6, NSX_F_PF = 7, NSX_F_UFNP = 8, NSX_F_VFNP = 9,   &
           NSX_F_WFNP = 10, NSX_F_PFNP = 11, NSX_F_RHSUF = 12,           &
           NSX_F_RHSVF = 13, NSX_F_RHSWF = 14, NSX_F_EDDY = 15,          &
           NSX_F_FORCE = 16,                                             &
           NSX_F_UFV = 32, NSX_F_VFV = 33, NSX_F_WFV = 34,               &
           NSX_F_PFV = 35, NSX_F_U1FACE = 48, NSX_F_U2FACE = 49,         &
           NSX_F_U3FACE = 50, NSX_F_UTFACE = 51, NSX_F_UBFACE = 52,      &
           NSX_F_CELL_SCRATCH = 64, NSX_F_VERT_SCRATCH = 96,             &
           NSX_F_FACE_SCRATCH = 112
!     solver outcome (enum nsx_solve_exit)
      integer(c_int), parameter :: NSX_SOLVE_CONVERGED = 0,             &
           NSX_SOLVE_NO_SOLVE = 1, NSX_SOLVE_BREAKDOWN = 2,              &
           NSX_SOLVE_MAX_ITERS = 3
!     precision and refinement (enum nsx_precision, enum nsx_refine)
      integer(c_int), parameter :: NSX_PREC_FP64 = 0, NSX_PREC_MIXED = 1,&
           NSX_PREC_FP64_CG = 2, NSX_PREC_MIXED_FLUX32 = 3
      integer(c_int), parameter :: NSX_REFINE_CG = 3,                   &
           NSX_REFINE_CGSR = 4, NSX_REFINE_CGSR_BLOCK = 5,               &
           NSX_REFINE_CGSR_LINE = 6, NSX_REFINE_MODAL = 7
!     initial guess of nsx_step's pressure solves (enum nsx_guess)
      integer(c_int), parameter :: NSX_GUESS_PREVIOUS = 0,              &
           NSX_GUESS_EXTRAPOLATE = 1
!     accumulators of the channel statistics (enum nsx_sm, module stats)
      integer(c_int), parameter :: NSX_SM_UF = 0, NSX_SM_VF = 1,        &
           NSX_SM_WF = 2, NSX_SM_U2F = 3, NSX_SM_V2F = 4, NSX_SM_W2F = 5,&
           NSX_SM_UVF = 6, NSX_SM_VWF = 7, NSX_SM_WUF = 8, NSX_SM_PF = 9,&
           NSX_SM_P2F = 10, NSX_SM_EF = 11

      type, bind(C) :: nsx_cfl_info
        real(c_double) :: umax, vmax, wmax, cfl, aveu, avev, avew
        real(c_double) :: flux_total, flux_max, flux_bc, flux_bot, flux_top, flux_side
      end type nsx_cfl_info

      type, bind(C) :: nsx_solve_info
        integer(c_int) :: iters, exit_code
        real(c_double) :: e1norm, e2norm
        integer(c_int) :: n_solves, iters_total, outer_passes, iters_fp32
      end type nsx_solve_info

      interface
        integer(c_int) function nsx_create(n_cell0, n_cell, n_vert, nz, &
             ctx) bind(C, name="nsx_create")
          import :: c_int, c_ptr
          integer(c_int), value :: n_cell0, n_cell, n_vert, nz
          type(c_ptr) :: ctx
        end function
        integer(c_int) function nsx_destroy(ctx) bind(C, name="nsx_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        type(c_ptr) function nsx_last_error(ctx)                         &
             bind(C, name="nsx_last_error")
          import :: c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function nsx_device(ctx) bind(C, name="nsx_device")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function nsx_set_mesh(ctx, No_cp, No_vp, nbe,     &
             nbev) bind(C, name="nsx_set_mesh")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), intent(in) :: No_cp(*), No_vp(*), nbe(*), nbev(*)
        end function
        integer(c_int) function nsx_set_geometry(ctx, areaCell, dlVV,    &
             dlCC, dxCC, dyCC, xnn, ynn, dle1, dle2, doe, sum_xc2,        &
             sum_yc2, sum_xcyc, TotalVolume)                              &
             bind(C, name="nsx_set_geometry")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: areaCell(*), dlVV(*), dlCC(*),   &
               dxCC(*), dyCC(*), xnn(*), ynn(*), dle1(*), dle2(*),        &
               doe(*), sum_xc2(*), sum_yc2(*), sum_xcyc(*)
          real(c_double), value :: TotalVolume
        end function
        integer(c_int) function nsx_set_sigma(ctx, sig, dsig, sigv,      &
             dsigv, sigIni, sigFin) bind(C, name="nsx_set_sigma")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: sig(*), dsig(*), sigv(*), dsigv(*)
          real(c_double), value :: sigIni, sigFin
        end function
        integer(c_int) function nsx_set_vertex_weights(ctx, surrounding, &
             weight, Dimsurrounding, dlVsum, vertex_pair, vertex_edge)    &
             bind(C, name="nsx_set_vertex_weights")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), intent(in) :: surrounding(*), Dimsurrounding(*), &
               vertex_pair(*), vertex_edge(*)
          real(c_double), intent(in) :: weight(*), dlVsum(*)
        end function
        integer(c_int) function nsx_set_bc(ctx, bcuc, bcvc, bcwc, bcpc,  &
             bcuv, bcvv, bcwv, bcpv, bctop, bcbot, TopBC, BotBC)          &
             bind(C, name="nsx_set_bc")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), intent(in) :: bcuc(*), bcvc(*), bcwc(*),        &
               bcpc(*), bcuv(*), bcvv(*), bcwv(*), bcpv(*)
          integer(c_int), value :: bctop, bcbot, TopBC, BotBC
        end function
        integer(c_int) function nsx_set_ppe_matrix(ctx, AmC0, AmC1,      &
             AmC2, AmC3, AmCT, AmCB) bind(C, name="nsx_set_ppe_matrix")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: AmC0(*), AmC1(*), AmC2(*),       &
               AmC3(*), AmCT(*), AmCB(*)
        end function
        integer(c_int) function nsx_set_les(ctx, walld_level)            &
             bind(C, name="nsx_set_les")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: walld_level(*)
        end function
        integer(c_int) function nsx_set_params(ctx, dt, Re, bdx, bdy,    &
             bdz, eps, MaxIters, FinRK, Iturbu)                           &
             bind(C, name="nsx_set_params")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: dt, Re, bdx, bdy, bdz, eps
          integer(c_int), value :: MaxIters, FinRK, Iturbu
        end function
        integer(c_int) function nsx_commit(ctx) bind(C, name="nsx_commit")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function nsx_put_field(ctx, field, host)          &
             bind(C, name="nsx_put_field")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: field
          real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function nsx_get_field(ctx, field, host,          &
             fill_ghost_rows) bind(C, name="nsx_get_field")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: field, fill_ghost_rows
          real(c_double), intent(inout) :: host(*)
        end function
        integer(c_int) function nsx_copy_field(ctx, dst, src)            &
             bind(C, name="nsx_copy_field")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: dst, src
        end function
        integer(c_int) function nsx_init_state(ctx)                      &
             bind(C, name="nsx_init_state")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function nsx_hydro(ctx, flag_ab, info)            &
             bind(C, name="nsx_hydro")
          import :: c_int, c_ptr, nsx_solve_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: flag_ab
          type(nsx_solve_info) :: info
        end function
        integer(c_int) function nsx_step(ctx, nsteps, precision, info)   &
             bind(C, name="nsx_step")
          import :: c_int, c_ptr, nsx_solve_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: nsteps, precision
          type(nsx_solve_info) :: info
        end function
        integer(c_int) function nsx_solve_info_get(ctx, info)            &
             bind(C, name="nsx_solve_info_get")
          import :: c_int, c_ptr, nsx_solve_info
          type(c_ptr), value :: ctx
          type(nsx_solve_info) :: info
        end function
        integer(c_int) function nsx_bcglobalvc(ctx, u, v, w, uv, vv, wv, &
             tagdir) bind(C, name="nsx_bcglobalvc")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: u, v, w, uv, vv, wv, tagdir
        end function
        integer(c_int) function nsx_bcglobalp(ctx, p, pv, tagdir)        &
             bind(C, name="nsx_bcglobalp")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: p, pv, tagdir
        end function
        integer(c_int) function nsx_grad_lsm(ctx, dfdx, dfdy, dfds, f)   &
             bind(C, name="nsx_grad_lsm")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: dfdx, dfdy, dfds, f
        end function
        integer(c_int) function nsx_calcul_face(ctx, f1, f2, f3, fT, fB, &
             fun, funv) bind(C, name="nsx_calcul_face")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: f1, f2, f3, fT, fB, fun, funv
        end function
        integer(c_int) function nsx_sgles(ctx, u, v, w)                  &
             bind(C, name="nsx_sgles")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: u, v, w
        end function
        integer(c_int) function nsx_advdiff(ctx, rhsu, rhsv, rhsw, u, v, &
             w, eddy) bind(C, name="nsx_advdiff")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: rhsu, rhsv, rhsw, u, v, w, eddy
        end function
        integer(c_int) function nsx_source3d(ctx, src, u, v, w)          &
             bind(C, name="nsx_source3d")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: src, u, v, w
        end function
        integer(c_int) function nsx_update_rhsp(ctx, rhs, p, pv)         &
             bind(C, name="nsx_update_rhsp")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: rhs, p, pv
        end function
        integer(c_int) function nsx_bicgstab(ctx, phi, phiv, rhs, info)  &
             bind(C, name="nsx_bicgstab")
          import :: c_int, c_ptr, nsx_solve_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: phi, phiv, rhs
          type(nsx_solve_info) :: info
        end function
        integer(c_int) function nsx_bicgstab_solve(ctx, phi, rhs, info)  &
             bind(C, name="nsx_bicgstab_solve")
          import :: c_int, c_ptr, nsx_solve_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: phi, rhs
          type(nsx_solve_info) :: info
        end function
        integer(c_int) function nsx_advdiff_gam(ctx, rhsu, rhsv, rhsw,   &
             u, v, w, gamx, gamy, gamz) bind(C, name="nsx_advdiff_gam")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: rhsu, rhsv, rhsw, u, v, w, gamx, gamy, gamz
        end function
        integer(c_int) function nsx_interpolate_u(ctx, f1, f2, f3, fT,   &
             fB, u, v, w) bind(C, name="nsx_interpolate_u")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: f1, f2, f3, fT, fB, u, v, w
        end function
        integer(c_int) function nsx_newu3d(ctx, p, pv)                   &
             bind(C, name="nsx_newu3d")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: p, pv
        end function
        integer(c_int) function nsx_bicgstab_adjoint(ctx, lam, g, info)  &
             bind(C, name="nsx_bicgstab_adjoint")
          import :: c_int, c_ptr, nsx_solve_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: lam, g
          type(nsx_solve_info) :: info
        end function
        integer(c_int) function nsx_set_body_force(ctx, on)              &
             bind(C, name="nsx_set_body_force")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: on
        end function
        integer(c_int) function nsx_wall_shear_run(ctx, nsteps, every,   &
             J, grad) bind(C, name="nsx_wall_shear_run")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: nsteps, every, grad
          real(c_double) :: J
        end function
        integer(c_int) function nsx_wall_shear_tangent(ctx, nsteps, dir, &
             J, dJ) bind(C, name="nsx_wall_shear_tangent")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: nsteps, dir
          real(c_double) :: J, dJ
        end function
        integer(c_int) function nsx_step_tangent(ctx, tin, tout, dforce) &
             bind(C, name="nsx_step_tangent")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int) :: tin(9), tout(9)
          integer(c_int), value :: dforce
        end function
        integer(c_int) function nsx_step_adjoint(ctx, bin, bout, gforce) &
             bind(C, name="nsx_step_adjoint")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int) :: bin(9), bout(9)
          integer(c_int), value :: gforce
        end function
        integer(c_int) function nsx_glob_cfl(ctx, u, v, w, info)         &
             bind(C, name="nsx_glob_cfl")
          import :: c_int, c_ptr, nsx_cfl_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: u, v, w
          type(nsx_cfl_info) :: info
        end function
        integer(c_int) function nsx_glob_cfl_noalloc(ctx, u, v, w, info) &
             bind(C, name="nsx_glob_cfl_noalloc")
          import :: c_int, c_ptr, nsx_cfl_info
          type(c_ptr), value :: ctx
          integer(c_int), value :: u, v, w
          type(nsx_cfl_info) :: info
        end function
        integer(c_int) function nsx_sm_init(ctx)                         &
             bind(C, name="nsx_sm_init")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function nsx_glob_sample(ctx, u, v, w, p)         &
             bind(C, name="nsx_glob_sample")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: u, v, w, p
        end function
        integer(c_int) function nsx_sm_count(ctx, n)                     &
             bind(C, name="nsx_sm_count")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int) :: n
        end function
        integer(c_int) function nsx_sm_get(ctx, which, host)             &
             bind(C, name="nsx_sm_get")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: which
          real(c_double) :: host(*)
        end function
        integer(c_int) function nsx_sm_planes(ctx, planes, ns)           &
             bind(C, name="nsx_sm_planes")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double) :: planes(*)
          integer(c_int) :: ns
        end function
        integer(c_int) function nsx_glob_put_plane_mm(ctx, dir)          &
             bind(C, name="nsx_glob_put_plane_mm")
          import :: c_int, c_ptr, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), dimension(*) :: dir
        end function
        integer(c_int) function nsx_write_plane_stats(planes, nz,        &
             n_cell0, ns, re, dir) bind(C, name="nsx_write_plane_stats")
          import :: c_int, c_double, c_char
          real(c_double) :: planes(*)
          integer(c_int), value :: nz, n_cell0, ns
          real(c_double), value :: re
          character(kind=c_char), dimension(*) :: dir
        end function
        integer(c_int) function nsx_format_real(x, kind, w, d, out,      &
             outlen) bind(C, name="nsx_format_real")
          import :: c_int, c_double, c_char
          real(c_double), value :: x
          integer(c_int), value :: kind, w, d, outlen
          character(kind=c_char), dimension(*) :: out
        end function
        integer(c_int) function nsx_savetec_vertex(ctx, dir, counter,    &
             time, xvt, yvt, zvt) bind(C, name="nsx_savetec_vertex")
          import :: c_int, c_ptr, c_double, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter
          real(c_double), value :: time
          real(c_double) :: xvt(*), yvt(*), zvt(*)
        end function
        integer(c_int) function nsx_savetec_center(ctx, dir, counter,    &
             time, xvt, yvt, zvt, u, v, w, p)                           &
             bind(C, name="nsx_savetec_center")
          import :: c_int, c_ptr, c_double, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter, u, v, w, p
          real(c_double), value :: time
          real(c_double) :: xvt(*), yvt(*), zvt(*)
        end function
        integer(c_int) function nsx_write_tec_vertex(dir, counter, time, &
             n_cell0, n_vert, nz, no_vp, xvt, yvt, zvt, ufv, vfv, wfv,   &
             pfv) bind(C, name="nsx_write_tec_vertex")
          import :: c_int, c_double, c_char
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter, n_cell0, n_vert, nz
          real(c_double), value :: time
          integer(c_int) :: no_vp(*)
          real(c_double) :: xvt(*), yvt(*), zvt(*), ufv(*), vfv(*), wfv(*), pfv(*)
        end function
        integer(c_int) function nsx_write_tec_center(dir, counter, time, &
             n_cell0, n_cell, n_vert, nz, no_vp, xvt, yvt, zvt, u, v, w, &
             p) bind(C, name="nsx_write_tec_center")
          import :: c_int, c_double, c_char
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter, n_cell0, n_cell, n_vert, nz
          real(c_double), value :: time
          integer(c_int) :: no_vp(*)
          real(c_double) :: xvt(*), yvt(*), zvt(*), u(*), v(*), w(*), p(*)
        end function
        integer(c_int) function nsx_save_paraview_vertex(ctx, dir,       &
             counter, xvt, yvt, zvt) bind(C, name="nsx_save_paraview_vertex")
          import :: c_int, c_ptr, c_double, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter
          real(c_double) :: xvt(*), yvt(*), zvt(*)
        end function
        integer(c_int) function nsx_savetec_vc(ctx, dir, counter, time,  &
             xvt, yvt, zvt, xct, yct, zct, u, v, w, p)                  &
             bind(C, name="nsx_savetec_vc")
          import :: c_int, c_ptr, c_double, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter, u, v, w, p
          real(c_double), value :: time
          real(c_double) :: xvt(*), yvt(*), zvt(*), xct(*), yct(*), zct(*)
        end function
        integer(c_int) function nsx_write_vtk_vertex(dir, counter,       &
             n_cell0, n_vert, nz, no_vp, xvt, yvt, zvt, ufv, vfv, wfv,   &
             usv, vsv, wsv) bind(C, name="nsx_write_vtk_vertex")
          import :: c_int, c_double, c_char
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter, n_cell0, n_vert, nz
          integer(c_int) :: no_vp(*)
          real(c_double) :: xvt(*), yvt(*), zvt(*), ufv(*), vfv(*), wfv(*)
          real(c_double) :: usv(*), vsv(*), wsv(*)
        end function
        integer(c_int) function nsx_write_tec_vc(dir, counter, time,     &
             n_cell0, n_cell, n_vert, nz, no_vp, xvt, yvt, zvt, xct, yct,&
             zct, ufv, vfv, wfv, pfv, u, v, w, p)                       &
             bind(C, name="nsx_write_tec_vc")
          import :: c_int, c_double, c_char
          character(kind=c_char), dimension(*) :: dir
          integer(c_int), value :: counter, n_cell0, n_cell, n_vert, nz
          real(c_double), value :: time
          integer(c_int) :: no_vp(*)
          real(c_double) :: xvt(*), yvt(*), zvt(*), xct(*), yct(*), zct(*)
          real(c_double) :: ufv(*), vfv(*), wfv(*), pfv(*), u(*), v(*), w(*), p(*)
        end function
        integer(c_int) function nsx_set_precision(ctx, precision, refine,&
             tol, max_passes) bind(C, name="nsx_set_precision")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: precision, refine, max_passes
          real(c_double), value :: tol
        end function
        integer(c_int) function nsx_set_initial_guess(ctx, mode)         &
             bind(C, name="nsx_set_initial_guess")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: mode
        end function
        integer(c_int) function nsx_timing_enable(ctx, on)               &
             bind(C, name="nsx_timing_enable")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: on
        end function
        integer(c_int) function nsx_timing_get(ctx, name, total_ms,      &
             launches) bind(C, name="nsx_timing_get")
          import :: c_int, c_ptr, c_char, c_double, c_long_long
          type(c_ptr), value :: ctx
          character(kind=c_char), intent(in) :: name(*)
          real(c_double) :: total_ms
          integer(c_long_long) :: launches
        end function
        integer(c_int) function nsx_timing_reset(ctx)                    &
             bind(C, name="nsx_timing_reset")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
      end interface

      contains

!     the text of nsx_last_error as a Fortran string
      function nsx_error_text(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: n
        msg = ' '
        p = nsx_last_error(ctx)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [512])
        do n = 1, 512
          if (s(n) == c_null_char) exit
          msg(n:n) = s(n)
        enddo
      end function nsx_error_text

      end module nsx_iface
