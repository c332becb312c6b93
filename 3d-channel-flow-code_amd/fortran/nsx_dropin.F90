!  nsx_dropin.F90 -- Fortran host side of the drop-in: the reference's hot-path
!  subroutines under their own names and argument lists, over the C ABI.
!
!  module nsx_dropin holds the device context and the sizes of COMMON /coo/
!  (N_CELL0, N_CELL, N_VERT, NZ; common.mpf:323-325).  nsx_dropin_setup hands
!  the reference's set-up arrays (module geometry after calcul_geometry and
!  PDE_center, module les, the run parameters) to the device once.  The
!  external subroutines below replace the reference routines of the same name:
!  they upload their array arguments, run the routine on the device and
!  download the results, so a caller that passes host arrays sees the
!  reference's interface.  Module state of the reference that the device keeps
!  (U1FACE..UBFACE, rhsuf/rhsvf/rhswf, eddy) stays resident between calls.
!  For the time loop itself nsx_dropin_step runs whole steps on the resident
!  state without host transfers.
!
!  Errors follow the reference: a fatal error prints and stops; a solver that
!  does not converge prints and returns (bicgstab.F90:562-566).
      module nsx_dropin
      use iso_c_binding
      use nsx_iface
      implicit none
      type(c_ptr), save :: nsx_ctx = c_null_ptr
      integer, save :: nsx_n_cell0 = 0, nsx_n_cell = 0, nsx_n_vert = 0, nsx_nz = 0
      type(nsx_solve_info), save :: nsx_last_info
!     checksum of the PPE matrix handed over at set-up (BICGSTAB_SOLVE)
      real(c_double), save :: nsx_ppe_sum = 0.d0

      contains

      subroutine nsx_dropin_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc < 0) then
          write(*,'(a,a,a,a)') ' nsx: ', where, ': ', trim(nsx_error_text(nsx_ctx))
          stop 'nsx: fatal error'
        else if (rc > 0) then
          write(*,'(a,a,a,i3)') ' nsx: ', where, ': solver outcome ', rc
        endif
      end subroutine nsx_dropin_check

!     returns 0, or the (negative) code of nsx_create when no device is usable
      integer function nsx_dropin_open(N_CELL0, N_CELL, N_VERT, NZ)
        integer, intent(in) :: N_CELL0, N_CELL, N_VERT, NZ
        nsx_n_cell0 = N_CELL0
        nsx_n_cell = N_CELL
        nsx_n_vert = N_VERT
        nsx_nz = NZ
        nsx_dropin_open = nsx_create(N_CELL0, N_CELL, N_VERT, NZ, nsx_ctx)
      end function nsx_dropin_open

      subroutine nsx_dropin_setup(No_cp, No_vp, nbe, nbev,               &
           areaCell, dlVV, dlCC, dxCC, dyCC, xnn, ynn, dle1, dle2, doe,   &
           sum_xc2, sum_yc2, sum_xcyc, TotalVolume,                       &
           sig, dsig, sigv, dsigv, sigIni, sigFin,                        &
           surrounding, weight, Dimsurrounding, dlVsum, vertex_pair,      &
           vertex_edge, bcuc, bcvc, bcwc, bcpc, bcuv, bcvv, bcwv, bcpv,   &
           bctop, bcbot, TopBC, BotBC, AmC0, AmC1, AmC2, AmC3, AmCT, AmCB,&
           walld_level, dt, Re, bdx, bdy, bdz, eps, MaxIters, FinRK, Iturbu)
        integer(c_int), intent(in) :: No_cp(*), No_vp(*), nbe(*), nbev(*)
        real(c_double), intent(in) :: areaCell(*), dlVV(*), dlCC(*),     &
             dxCC(*), dyCC(*), xnn(*), ynn(*), dle1(*), dle2(*), doe(*),  &
             sum_xc2(*), sum_yc2(*), sum_xcyc(*)
        real(c_double), intent(in) :: TotalVolume, sigIni, sigFin
        real(c_double), intent(in) :: sig(*), dsig(*), sigv(*), dsigv(*)
        integer(c_int), intent(in) :: surrounding(*), Dimsurrounding(*), &
             vertex_pair(*), vertex_edge(*)
        real(c_double), intent(in) :: weight(*), dlVsum(*)
        integer(c_int), intent(in) :: bcuc(*), bcvc(*), bcwc(*), bcpc(*),&
             bcuv(*), bcvv(*), bcwv(*), bcpv(*)
        integer(c_int), intent(in) :: bctop, bcbot, TopBC, BotBC
        real(c_double), intent(in) :: AmC0(*), AmC1(*), AmC2(*), AmC3(*),&
             AmCT(*), AmCB(*), walld_level(*)
        real(c_double), intent(in) :: dt, Re, bdx, bdy, bdz, eps
        integer(c_int), intent(in) :: MaxIters, FinRK, Iturbu
        call nsx_dropin_check(nsx_set_mesh(nsx_ctx, No_cp, No_vp, nbe,   &
             nbev), 'nsx_set_mesh')
        call nsx_dropin_check(nsx_set_geometry(nsx_ctx, areaCell, dlVV,  &
             dlCC, dxCC, dyCC, xnn, ynn, dle1, dle2, doe, sum_xc2,        &
             sum_yc2, sum_xcyc, TotalVolume), 'nsx_set_geometry')
        call nsx_dropin_check(nsx_set_sigma(nsx_ctx, sig, dsig, sigv,    &
             dsigv, sigIni, sigFin), 'nsx_set_sigma')
        call nsx_dropin_check(nsx_set_vertex_weights(nsx_ctx,            &
             surrounding, weight, Dimsurrounding, dlVsum, vertex_pair,    &
             vertex_edge), 'nsx_set_vertex_weights')
        call nsx_dropin_check(nsx_set_bc(nsx_ctx, bcuc, bcvc, bcwc, bcpc,&
             bcuv, bcvv, bcwv, bcpv, bctop, bcbot, TopBC, BotBC),         &
             'nsx_set_bc')
        call nsx_dropin_check(nsx_set_ppe_matrix(nsx_ctx, AmC0, AmC1,    &
             AmC2, AmC3, AmCT, AmCB), 'nsx_set_ppe_matrix')
        nsx_ppe_sum = nsx_matrix_sum(AmC0, AmC1, AmC2, AmC3, AmCT, AmCB)
        call nsx_dropin_check(nsx_set_les(nsx_ctx, walld_level),         &
             'nsx_set_les')
        call nsx_dropin_check(nsx_set_params(nsx_ctx, dt, Re, bdx, bdy,  &
             bdz, eps, MaxIters, FinRK, Iturbu), 'nsx_set_params')
        call nsx_dropin_check(nsx_commit(nsx_ctx), 'nsx_commit')
      end subroutine nsx_dropin_setup

!     the resident time loop: nsteps of nsmp3D.F90:404-564 on the device
      subroutine nsx_dropin_step(nsteps, precision)
        integer, intent(in) :: nsteps, precision
        call nsx_dropin_check(nsx_step(nsx_ctx, nsteps, precision,       &
             nsx_last_info), 'nsx_step')
      end subroutine nsx_dropin_step

      subroutine nsx_dropin_put(field, a)
        integer(c_int), intent(in) :: field
        real(c_double), intent(in) :: a(*)
        call nsx_dropin_check(nsx_put_field(nsx_ctx, field, a), 'nsx_put_field')
      end subroutine nsx_dropin_put

      subroutine nsx_dropin_get(field, a)
        integer(c_int), intent(in) :: field
        real(c_double), intent(inout) :: a(*)
        call nsx_dropin_check(nsx_get_field(nsx_ctx, field, a, 1), 'nsx_get_field')
      end subroutine nsx_dropin_get

!     a field whose reference routine writes the interior cells only
!     (rows 1..N_CELL0): the ghost rows of the caller's array are kept
      subroutine nsx_dropin_get_interior(field, a)
        integer(c_int), intent(in) :: field
        real(c_double), intent(inout) :: a(nsx_n_cell,nsx_nz)
        real(c_double), allocatable :: t(:,:)
        allocate(t(nsx_n_cell,nsx_nz))
        call nsx_dropin_check(nsx_get_field(nsx_ctx, field, t, 0), 'nsx_get_field')
        a(1:nsx_n_cell0,:) = t(1:nsx_n_cell0,:)
        deallocate(t)
      end subroutine nsx_dropin_get_interior

!     the coordinates of the cell centres and vertices of LoopTimeInitial
!     and LoopTimeUpdate: x, y of the column, z = sig Hpr - h
      subroutine nsx_dropin_coords(xct, yct, zct, xvt, yvt, zvt, xc, yc, &
           sig, Hpr, h, xv, yv, sigv, Hprv, hv)
        real*8, dimension(:,:) :: xct, yct, zct, xvt, yvt, zvt
        real*8, dimension(:) :: xc, yc, sig, Hpr, h, xv, yv, sigv, Hprv, hv
        integer :: i, k, nv
        do i = 1, nsx_n_cell
          do k = 1, nsx_nz
            xct(i,k) = xc(i)
            yct(i,k) = yc(i)
            zct(i,k) = sig(k)*Hpr(i)-h(i)
          enddo
        enddo
        do nv = 1, nsx_n_vert
          do k = 1, nsx_nz-1
            xvt(nv,k) = xv(nv)
            yvt(nv,k) = yv(nv)
            zvt(nv,k) = sigv(k)*Hprv(nv)-hv(nv)
          enddo
        enddo
      end subroutine nsx_dropin_coords

!     a checksum of the six PPE coefficient arrays (N_CELL0,NZ), each entry
!     weighted by its position, summed in a fixed order
      real(c_double) function nsx_matrix_sum(A0, A1, A2, A3, AT, AB)
        real(c_double), dimension(nsx_n_cell0*nsx_nz) :: A0, A1, A2, A3, AT, AB
        integer :: n
        real(c_double) :: s
        s = 0.d0
        do n = 1, nsx_n_cell0*nsx_nz
          s = s + dble(mod(n,7)+1)*A0(n) + dble(mod(n,11)+1)*A1(n)        &
                + dble(mod(n,13)+1)*A2(n) + dble(mod(n,17)+1)*A3(n)       &
                + dble(mod(n,19)+1)*AT(n) + dble(mod(n,23)+1)*AB(n)
        enddo
        nsx_matrix_sum = s
      end function nsx_matrix_sum

      subroutine nsx_dropin_close()
        integer(c_int) :: rc
        if (c_associated(nsx_ctx)) rc = nsx_destroy(nsx_ctx)
        nsx_ctx = c_null_ptr
      end subroutine nsx_dropin_close

      end module nsx_dropin

!-----------------------------------------------------------------------
!  bicgstab.F90:8-168 -- solution of the pressure Poisson equation
      SUBROUTINE bicgstab(phi,phiv,rhs,                                  &
                          xc,yc,sig,dsig,No_cp,nbe,                      &
                          xv,yv,sigv,dsigv,No_vp,nbev,                   &
                          Hpr,h,etan,                                    &
                          Hprv,hv,etav)
      use nsx_dropin
      implicit none
      real*8 :: phi(nsx_n_cell,nsx_nz), phiv(nsx_n_vert,nsx_nz-1)
      real*8 :: rhs(nsx_n_cell,nsx_nz)
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      real*8 :: Hpr(nsx_n_cell), h(nsx_n_cell), etan(nsx_n_cell)
      real*8 :: Hprv(nsx_n_vert), hv(nsx_n_vert), etav(nsx_n_vert)
      integer(c_int) :: rc
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phi)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, rhs)
      rc = nsx_bicgstab(nsx_ctx, NSX_F_CELL_SCRATCH, NSX_F_VERT_SCRATCH, &
                        NSX_F_CELL_SCRATCH + 1, nsx_last_info)
      call nsx_dropin_check(rc, 'bicgstab')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, phi)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH, phiv)
      END SUBROUTINE bicgstab

!-----------------------------------------------------------------------
!  hydro.F90:8-501 -- one projection stage
      SUBROUTINE hydro(ufnp,ufn,uf,ufv,                                  &
                       vfnp,vfn,vf,vfv,                                  &
                       wfnp,wfn,wf,wfv,                                  &
                       pfnp,pfn,pf,pfv,                                  &
                       rhof,viscof,                                      &
                       rhofv,viscofv,                                    &
                       xc,yc,sig,dsig,No_cp,nbe,                         &
                       xv,yv,sigv,dsigv,No_vp,nbev,                      &
                       Hpr,h,eta,etan,                                   &
                       Hprv,hv,etav,                                     &
                       flag_ab)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: ufnp, ufn, uf, vfnp, vfn, vf, &
           wfnp, wfn, wf, pfnp, pfn, pf, rhof, viscof
      real*8, dimension(nsx_n_vert,nsx_nz-1) :: ufv, vfv, wfv, pfv, rhofv, viscofv
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      real*8 :: Hpr(nsx_n_cell), h(nsx_n_cell), eta(nsx_n_cell), etan(nsx_n_cell)
      real*8 :: Hprv(nsx_n_vert), hv(nsx_n_vert), etav(nsx_n_vert)
      integer :: flag_ab
      integer(c_int) :: rc
!     stage input (uf..pf) and time level n; the module state of the reference
!     (faces, AB2 history, eddy) is the device's resident copy
      call nsx_dropin_put(NSX_F_UF, uf)
      call nsx_dropin_put(NSX_F_VF, vf)
      call nsx_dropin_put(NSX_F_WF, wf)
      call nsx_dropin_put(NSX_F_PF, pf)
      call nsx_dropin_put(NSX_F_UFN, ufn)
      call nsx_dropin_put(NSX_F_VFN, vfn)
      call nsx_dropin_put(NSX_F_WFN, wfn)
      call nsx_dropin_put(NSX_F_PFN, pfn)
      rc = nsx_hydro(nsx_ctx, flag_ab, nsx_last_info)
      call nsx_dropin_check(rc, 'hydro')
      call nsx_dropin_get(NSX_F_UFNP, ufnp)
      call nsx_dropin_get(NSX_F_VFNP, vfnp)
      call nsx_dropin_get(NSX_F_WFNP, wfnp)
      call nsx_dropin_get(NSX_F_PFNP, pfnp)
      call nsx_dropin_get(NSX_F_UFV, ufv)
      call nsx_dropin_get(NSX_F_VFV, vfv)
      call nsx_dropin_get(NSX_F_WFV, wfv)
      call nsx_dropin_get(NSX_F_PFV, pfv)
      END SUBROUTINE hydro

!-----------------------------------------------------------------------
!  gradientLSM.F90:8-211 -- least-squares cell gradients
      SUBROUTINE grandientLSM(dfundx,dfundy,dfundsig,                    &
                              fun,xc,yc,sig,dsig,No_cp,nbe)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: dfundx, dfundy, dfundsig, fun
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 3, fun)
      call nsx_dropin_check(nsx_grad_lsm(nsx_ctx, NSX_F_CELL_SCRATCH,    &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2,                &
           NSX_F_CELL_SCRATCH + 3), 'grandientLSM')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, dfundx)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, dfundy)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, dfundsig)
      END SUBROUTINE grandientLSM

!-----------------------------------------------------------------------
!  The remaining routines of src/AdvectionDiffusion, src/Gradient, src/LES,
!  src/Interpolation, src/BoundaryConditions, the pressure solver and
!  src/LoopUpdate under the reference's names and argument lists.  Each puts
!  its array arguments into the device's scratch fields, runs the routine on
!  the device and gets the arrays the reference routine writes (the input
!  fields too where the reference fills their ghost cells in place).  The
!  geometry arguments (xc ... nbev) are the set-up's and are not read again.

!  BC.F90:202-295 -- ghost cells of u, v, w and their vertex values
      SUBROUTINE BCglobalVC(phiuc,phivc,phiwc,                           &
                            phiuv,phivv,phiwv,                           &
                            xc,yc,sig,dsig,No_cp,nbe,                    &
                            xv,yv,sigv,dsigv,No_vp,nbev,tagdir)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: phiuc, phivc, phiwc
      real*8, dimension(nsx_n_vert,nsx_nz-1) :: phiuv, phivv, phiwv
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      integer :: tagdir
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phiuc)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, phivc)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 2, phiwc)
      call nsx_dropin_check(nsx_bcglobalvc(nsx_ctx, NSX_F_CELL_SCRATCH,  &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2,                &
           NSX_F_VERT_SCRATCH, NSX_F_VERT_SCRATCH + 1,                    &
           NSX_F_VERT_SCRATCH + 2, tagdir), 'BCglobalVC')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, phiuc)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, phivc)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, phiwc)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH, phiuv)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH + 1, phivv)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH + 2, phiwv)
      END SUBROUTINE BCglobalVC

!  BC.F90:299-363 -- ghost cells of p and its vertex values
      SUBROUTINE BCglobalP(phi,phiv,                                     &
                           xc,yc,sig,dsig,No_cp,nbe,                     &
                           xv,yv,sigv,dsigv,No_vp,nbev,tagdir)
      use nsx_dropin
      implicit none
      real*8 :: phi(nsx_n_cell,nsx_nz), phiv(nsx_n_vert,nsx_nz-1)
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      integer :: tagdir
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phi)
      call nsx_dropin_check(nsx_bcglobalp(nsx_ctx, NSX_F_CELL_SCRATCH,   &
           NSX_F_VERT_SCRATCH, tagdir), 'BCglobalP')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, phi)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH, phiv)
      END SUBROUTINE BCglobalP

!  gradientP.F90:8-150 -- the pressure gradient, Option 2 (grandientLSM)
      SUBROUTINE gradientP(fun,funv,                                     &
                           dfundx,dfundy,dfundsig,                       &
                           xc,yc,sig,dsig,No_cp,nbe,                     &
                           xv,yv,sigv,dsigv,No_vp,nbev)
      use nsx_dropin
      implicit none
      real*8 :: fun(nsx_n_cell,nsx_nz), funv(nsx_n_vert,nsx_nz-1)
      real*8, dimension(nsx_n_cell,nsx_nz) :: dfundx, dfundy, dfundsig
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 3, fun)
      call nsx_dropin_check(nsx_grad_lsm(nsx_ctx, NSX_F_CELL_SCRATCH,    &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2,                &
           NSX_F_CELL_SCRATCH + 3), 'gradientP')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, dfundx)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, dfundy)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, dfundsig)
      END SUBROUTINE gradientP

!  interpolationFace.F90:7-238 -- the five face values of fun (Kim-Choi,
!  whatever Option, as the reference's KeyKimChoi build)
      SUBROUTINE calcul_face(funf1,funf2,funf3,funfT,funfB,              &
                             fun,funv,                                   &
                             xc,yc,sig,dsig,No_cp,nbe,                   &
                             xv,yv,sigv,dsigv,No_vp,nbev,Option)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell0,nsx_nz) :: funf1, funf2, funf3, funfT, funfB
      real*8 :: fun(nsx_n_cell,nsx_nz), funv(nsx_n_vert,nsx_nz-1)
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      integer :: Option
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, fun)
      call nsx_dropin_put(NSX_F_VERT_SCRATCH, funv)
      call nsx_dropin_check(nsx_calcul_face(nsx_ctx, NSX_F_FACE_SCRATCH, &
           NSX_F_FACE_SCRATCH + 1, NSX_F_FACE_SCRATCH + 2,                &
           NSX_F_FACE_SCRATCH + 3, NSX_F_FACE_SCRATCH + 4,                &
           NSX_F_CELL_SCRATCH, NSX_F_VERT_SCRATCH), 'calcul_face')
      call nsx_dropin_get(NSX_F_FACE_SCRATCH, funf1)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 1, funf2)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 2, funf3)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 3, funfT)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 4, funfB)
      END SUBROUTINE calcul_face

!  interpolationU.F90:5-190 -- face-normal velocities of u, v, w
      SUBROUTINE interpolateU(un1,un2,un3,unT,unB,                       &
                              phiu,phiv,phiw,                            &
                              xc,yc,sig,dsig,No_cp,nbe,                  &
                              xv,yv,sigv,dsigv,No_vp,nbev)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell0,nsx_nz) :: un1, un2, un3, unT, unB
      real*8, dimension(nsx_n_cell,nsx_nz) :: phiu, phiv, phiw
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phiu)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, phiv)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 2, phiw)
      call nsx_dropin_check(nsx_interpolate_u(nsx_ctx,                   &
           NSX_F_FACE_SCRATCH, NSX_F_FACE_SCRATCH + 1,                    &
           NSX_F_FACE_SCRATCH + 2, NSX_F_FACE_SCRATCH + 3,                &
           NSX_F_FACE_SCRATCH + 4, NSX_F_CELL_SCRATCH,                    &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2), 'interpolateU')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, phiu)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, phiv)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, phiw)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH, un1)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 1, un2)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 2, un3)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 3, unT)
      call nsx_dropin_get(NSX_F_FACE_SCRATCH + 4, unB)
      END SUBROUTINE interpolateU

!  sgm_les.F90:24-275 -- the Smagorinsky eddy viscosity into module les'
!  eddy, which stays resident on the device (nsx_dropin_get(NSX_F_EDDY, .))
      subroutine sgles(phiu,phiv,phiw,                                   &
                       xc,yc,sig,dsig,No_cp,nbe,                         &
                       xv,yv,sigv,dsigv,No_vp,nbev,                      &
                       Hpr,h,eta,etan,                                   &
                       Hprv,hv,etav)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: phiu, phiv, phiw
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      real*8 :: Hpr(nsx_n_cell), h(nsx_n_cell), eta(nsx_n_cell), etan(nsx_n_cell)
      real*8 :: Hprv(nsx_n_vert), hv(nsx_n_vert), etav(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phiu)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, phiv)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 2, phiw)
      call nsx_dropin_check(nsx_sgles(nsx_ctx, NSX_F_CELL_SCRATCH,       &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2), 'sgles')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, phiu)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, phiv)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, phiw)
      end subroutine sgles

!  AdvDiffNS.F90:8-140 -- rhs? += Gam? x diffusion - advection, advecting
!  with the resident face velocities U1FACE..UBFACE
      SUBROUTINE AdvDiffNS(rhsu,rhsv,rhsw,                               &
                           funu,funv,funw,                               &
                           Gamx,Gamy,Gamz,                               &
                           xc,yc,sig,dsig,No_cp,nbe,                     &
                           xv,yv,sigv,dsigv,No_vp,nbev)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: rhsu, rhsv, rhsw, funu, funv, funw, Gamx, Gamy, Gamz
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, rhsu)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, rhsv)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 2, rhsw)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 3, funu)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 4, funv)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 5, funw)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 6, Gamx)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 7, Gamy)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 8, Gamz)
      call nsx_dropin_check(nsx_advdiff_gam(nsx_ctx, NSX_F_CELL_SCRATCH, &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2,                &
           NSX_F_CELL_SCRATCH + 3, NSX_F_CELL_SCRATCH + 4,                &
           NSX_F_CELL_SCRATCH + 5, NSX_F_CELL_SCRATCH + 6,                &
           NSX_F_CELL_SCRATCH + 7, NSX_F_CELL_SCRATCH + 8), 'AdvDiffNS')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, rhsu)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, rhsv)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, rhsw)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 3, funu)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 4, funv)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 5, funw)
      END SUBROUTINE AdvDiffNS

!  source3D.F90:8-174 -- face velocities of u*, v*, w* into the resident
!  U1FACE..UBFACE and their divergence into source
      SUBROUTINE Source3D(source,phiu,phiv,phiw,                         &
                          xc,yc,sig,dsig,No_cp,nbe,                      &
                          xv,yv,sigv,dsigv,No_vp,nbev)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: source, phiu, phiv, phiw
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, phiu)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 2, phiv)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 3, phiw)
      call nsx_dropin_check(nsx_source3d(nsx_ctx, NSX_F_CELL_SCRATCH,    &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2,                &
           NSX_F_CELL_SCRATCH + 3), 'Source3D')
      call nsx_dropin_get_interior(NSX_F_CELL_SCRATCH, source)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, phiu)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 2, phiv)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 3, phiw)
      END SUBROUTINE Source3D

!  source3D.F90:176-335 -- rhs -= the lagged Kim-Choi flux of p; the vertex
!  values of p (BCglobalP tagdir 2) into phiv
      SUBROUTINE update_rhsp(rhs,phi,phiv,dphidx,dphidy,dphids,          &
                             xc,yc,sig,dsig,No_cp,nbe,                   &
                             xv,yv,sigv,dsigv,No_vp,nbev)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: rhs, phi, dphidx, dphidy, dphids
      real*8 :: phiv(nsx_n_vert,nsx_nz-1)
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, rhs)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, phi)
      call nsx_dropin_check(nsx_update_rhsp(nsx_ctx, NSX_F_CELL_SCRATCH, &
           NSX_F_CELL_SCRATCH + 1, NSX_F_VERT_SCRATCH), 'update_rhsp')
      call nsx_dropin_get_interior(NSX_F_CELL_SCRATCH, rhs)
      call nsx_dropin_get(NSX_F_CELL_SCRATCH + 1, phi)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH, phiv)
      END SUBROUTINE update_rhsp

!  newu3D.F90:8-257 -- U1FACE..UBFACE -= dt grad_n p on the resident faces;
!  the vertex values of p (BCglobalP tagdir 2) into funv
      SUBROUTINE newu3D(fun,funv,flag,                                   &
                        dpdx,dpdy,dpds,                                  &
                        xc,yc,sig,dsig,No_cp,nbe,                        &
                        xv,yv,sigv,dsigv,No_vp,nbev)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: fun, dpdx, dpdy, dpds
      real*8 :: funv(nsx_n_vert,nsx_nz-1)
      integer :: flag
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      real*8 :: xv(nsx_n_vert), yv(nsx_n_vert), sigv(nsx_nz-1), dsigv(nsx_nz-1)
      integer :: No_vp(nsx_n_cell0,3), nbev(nsx_n_vert)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, fun)
      call nsx_dropin_check(nsx_newu3d(nsx_ctx, NSX_F_CELL_SCRATCH,      &
           NSX_F_VERT_SCRATCH), 'newu3D')
      call nsx_dropin_get(NSX_F_CELL_SCRATCH, fun)
      call nsx_dropin_get(NSX_F_VERT_SCRATCH, funv)
      END SUBROUTINE newu3D

!  bicgstab.F90:303-591 -- the BiCGSTAB iteration alone, on the set-up's PPE
!  matrix: the coefficients passed must be those (the matrix bicgstab copies
!  from AmC0..AmCB), which a checksum taken at nsx_dropin_setup confirms
      SUBROUTINE BICGSTAB_SOLVE(Am0,Am1,Am2,Am3,AmT,AmB,                 &
                                rhs,phi,                                 &
                                xc,yc,sig,dsig,No_cp,nbe,TagBC)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell0,nsx_nz) :: Am0, Am1, Am2, Am3, AmT, AmB
      real*8, dimension(nsx_n_cell,nsx_nz) :: rhs, phi
      real*8 :: xc(nsx_n_cell), yc(nsx_n_cell), sig(nsx_nz), dsig(nsx_nz)
      integer :: No_cp(nsx_n_cell,3), nbe(nsx_n_cell0)
      integer :: TagBC
      integer(c_int) :: rc
      if (nsx_matrix_sum(Am0, Am1, Am2, Am3, AmT, AmB) /= nsx_ppe_sum) then
        write(*,'(a)') ' nsx: BICGSTAB_SOLVE: the coefficients are not the PPE matrix of the set-up'
        stop 'nsx: fatal error'
      endif
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phi)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, rhs)
      rc = nsx_bicgstab_solve(nsx_ctx, NSX_F_CELL_SCRATCH,               &
                              NSX_F_CELL_SCRATCH + 1, nsx_last_info)
      call nsx_dropin_check(rc, 'BICGSTAB_SOLVE')
      call nsx_dropin_get_interior(NSX_F_CELL_SCRATCH, phi)
      END SUBROUTINE BICGSTAB_SOLVE

!  LoopTimeInitial.F90:8-259 -- the coordinates of the centres and vertices,
!  and level n+1 <- level n, on the caller's arrays and on the device's
!  resident time levels
      SUBROUTINE LoopTimeInitial(etanp,                                  &
                                 alphafnp,ufnp,vfnp,wfnp,pfnp,           &
                                 alphasnp,usnp,vsnp,wsnp,psnp,           &
                                 xct,yct,zct,                            &
                                 xvt,yvt,zvt,                            &
                                 etan,                                   &
                                 alphafn,ufn,vfn,wfn,pfn,                &
                                 alphasn,usn,vsn,wsn,psn,                &
                                 xc,yc,sig,Hpr,h,                        &
                                 xv,yv,sigv,Hprv,hv,                     &
                                 No_cp,nbe)
      use nsx_dropin
      implicit none
      real*8, dimension(:)   :: etanp, etan
      real*8, dimension(:,:) :: alphafnp, ufnp, vfnp, wfnp, pfnp
      real*8, dimension(:,:) :: alphasnp, usnp, vsnp, wsnp, psnp
      real*8, dimension(:,:) :: alphafn, ufn, vfn, wfn, pfn
      real*8, dimension(:,:) :: alphasn, usn, vsn, wsn, psn
      real*8, dimension(:,:) :: xct, yct, zct, xvt, yvt, zvt
      real*8, dimension(:)   :: xc, yc, sig, Hpr, h, xv, yv, sigv, Hprv, hv
      integer, dimension(:,:) :: No_cp
      integer, dimension(:)   :: nbe
      integer :: q
      call nsx_dropin_coords(xct, yct, zct, xvt, yvt, zvt, xc, yc, sig,  &
                             Hpr, h, xv, yv, sigv, Hprv, hv)
      etanp = etan
      alphafnp = alphafn
      alphasnp = alphasn
      ufnp = ufn
      vfnp = vfn
      wfnp = wfn
      usnp = usn
      vsnp = vsn
      wsnp = wsn
      pfnp = pfn
      psnp = psn
      do q = 0, 3
        call nsx_dropin_check(nsx_copy_field(nsx_ctx, NSX_F_UFNP + q,    &
             NSX_F_UFN + q), 'LoopTimeInitial')
      enddo
      END SUBROUTINE LoopTimeInitial

!  LoopTimeUpdate.F90:8-250 -- the rigid-surface depth (ChooseDomBox = 1:
!  h = 0, Hpr = 1), the coordinates, and level n <- level n+1, on the
!  caller's arrays and on the device's resident time levels
      SUBROUTINE LoopTimeUpdate(etan,                                    &
                                alphafn,ufn,vfn,wfn,pfn,                 &
                                alphasn,usn,vsn,wsn,psn,                 &
                                xct,yct,zct,                             &
                                xvt,yvt,zvt,                             &
                                etanp,                                   &
                                alphafnp,ufnp,vfnp,wfnp,pfnp,            &
                                alphasnp,usnp,vsnp,wsnp,psnp,            &
                                xc,yc,sig,Hpr,h,                         &
                                xv,yv,sigv,Hprv,hv,                      &
                                No_cp,nbe)
      use nsx_dropin
      implicit none
      real*8, dimension(:)   :: etanp, etan
      real*8, dimension(:,:) :: alphafnp, ufnp, vfnp, wfnp, pfnp
      real*8, dimension(:,:) :: alphasnp, usnp, vsnp, wsnp, psnp
      real*8, dimension(:,:) :: alphafn, ufn, vfn, wfn, pfn
      real*8, dimension(:,:) :: alphasn, usn, vsn, wsn, psn
      real*8, dimension(:,:) :: xct, yct, zct, xvt, yvt, zvt
      real*8, dimension(:)   :: xc, yc, sig, Hpr, h, xv, yv, sigv, Hprv, hv
      integer, dimension(:,:) :: No_cp
      integer, dimension(:)   :: nbe
      integer :: q
      h = 0.0d0
      Hpr = 1.0d0
      hv = 0.0d0
      Hprv = 1.0d0
      call nsx_dropin_coords(xct, yct, zct, xvt, yvt, zvt, xc, yc, sig,  &
                             Hpr, h, xv, yv, sigv, Hprv, hv)
      etan = etanp
      alphafn = alphafnp
      alphasn = alphasnp
      ufn = ufnp
      vfn = vfnp
      wfn = wfnp
      usn = usnp
      vsn = vsnp
      wsn = wsnp
      pfn = pfnp
      psn = psnp
      do q = 0, 3
        call nsx_dropin_check(nsx_copy_field(nsx_ctx, NSX_F_UFN + q,     &
             NSX_F_UFNP + q), 'LoopTimeUpdate')
      enddo
      END SUBROUTINE LoopTimeUpdate

!-----------------------------------------------------------------------
!  stats.F90:34-84 -- zero the channel statistics (device accumulators;
!  the three routines below replace stats.F90 as a whole)
      SUBROUTINE sm_init()
      use nsx_dropin
      implicit none
      call nsx_dropin_check(nsx_sm_init(nsx_ctx), 'sm_init')
      END SUBROUTINE sm_init

!-----------------------------------------------------------------------
!  stats.F90:91-197 -- one sample of u, v, w, p and of the eddy viscosity
!  (the device's resident copy of module les' eddy)
      SUBROUTINE glob_sample(phiu,phiv,phiw,phip)
      use nsx_dropin
      implicit none
      real*8, dimension(nsx_n_cell,nsx_nz) :: phiu, phiv, phiw, phip
      integer(c_int) :: n
      call nsx_dropin_put(NSX_F_CELL_SCRATCH, phiu)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 1, phiv)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 2, phiw)
      call nsx_dropin_put(NSX_F_CELL_SCRATCH + 3, phip)
      call nsx_dropin_check(nsx_glob_sample(nsx_ctx, NSX_F_CELL_SCRATCH, &
           NSX_F_CELL_SCRATCH + 1, NSX_F_CELL_SCRATCH + 2,                &
           NSX_F_CELL_SCRATCH + 3), 'glob_sample')
      call nsx_dropin_check(nsx_sm_count(nsx_ctx, n), 'glob_sample')
      if (mod(n,10) .eq. 0) then
         print*, '       =========================================='
         print*, '               STATS SAMPLED',n,'TIMES'
         print*, '       =========================================='
      endif
      END SUBROUTINE glob_sample

!-----------------------------------------------------------------------
!  stats.F90:204-391 -- plane-averaged statistics into ../output/Stats
      SUBROUTINE glob_put_plane_mm()
      use nsx_dropin
      implicit none
      integer(c_int) :: n
      call nsx_dropin_check(nsx_glob_put_plane_mm(nsx_ctx,               &
           '../output/Stats'//c_null_char), 'glob_put_plane_mm')
      call nsx_dropin_check(nsx_sm_count(nsx_ctx, n), 'glob_put_plane_mm')
      print*,'   --- Plane Averaged Stats printed! ---',n
      END SUBROUTINE glob_put_plane_mm

!-----------------------------------------------------------------------
!  The Tecplot snapshots of src/Save.  The reference declares both with
!  assumed-shape dummies (module_interfaces.F90:141-190), so callers pass
!  array descriptors and these take the same form.  SaveCounter and time
!  come from the caller's COMMON /integ/ and /times/, declared here with the
!  member types of common.mpf:221-228,310-346 (a COMMON block is a storage
!  sequence, so the members must keep the reference's order and kinds).
!  The files go to ../output/Serial, as the reference's serial build writes.

!  SavetecVertex.F90:185-235 -- V-NNNN.tec of the vertex fields
      SUBROUTINE SavetecVertex(alphafv,ufv,vfv,wfv,pfv,                  &
                               alphasv,usv,vsv,wsv,psv,rhosv,            &
                               xvt,yvt,zvt,No_vp)
      use nsx_dropin
      implicit none
      real*8, dimension(:,:) :: alphafv, ufv, vfv, wfv, pfv
      real*8, dimension(:,:) :: alphasv, usv, vsv, wsv, psv, rhosv
      real*8, dimension(:,:) :: xvt, yvt, zvt
      integer, dimension(:,:) :: No_vp
      real*8 :: tFin, tIni, dt, tInisave, dtsave, LastTimeSave, time
      integer :: tlimit, SaveCounter, Numictec, nc_ref, FinRK
      common/times/tFin,tIni,dt,tInisave,dtsave,LastTimeSave,time,tlimit
      common/integ/SaveCounter,Numictec,nc_ref,FinRK
      if (nsx_write_tec_vertex('../output/Serial'//c_null_char,          &
           SaveCounter, time, nsx_n_cell0, nsx_n_vert, nsx_nz, No_vp,    &
           xvt, yvt, zvt, ufv, vfv, wfv, pfv) /= 0) then
        write(*,'(a)') ' nsx: SavetecVertex: cannot write ../output/Serial'
        stop 'nsx: fatal error'
      endif
      END SUBROUTINE SavetecVertex

!  SavetecCenter.F90:185-235 -- C-NNNNN.tec of the cell fields
      SUBROUTINE SavetecCenter(alphafnp,ufnp,vfnp,wfnp,pfnp,             &
                               alphasnp,usnp,vsnp,wsnp,psnp,rhos,        &
                               xvt,yvt,zvt,No_vp)
      use nsx_dropin
      implicit none
      real*8, dimension(:,:) :: alphafnp, ufnp, vfnp, wfnp, pfnp
      real*8, dimension(:,:) :: alphasnp, usnp, vsnp, wsnp, psnp, rhos
      real*8, dimension(:,:) :: xvt, yvt, zvt
      integer, dimension(:,:) :: No_vp
      real*8 :: tFin, tIni, dt, tInisave, dtsave, LastTimeSave, time
      integer :: tlimit, SaveCounter, Numictec, nc_ref, FinRK
      common/times/tFin,tIni,dt,tInisave,dtsave,LastTimeSave,time,tlimit
      common/integ/SaveCounter,Numictec,nc_ref,FinRK
      if (nsx_write_tec_center('../output/Serial'//c_null_char,          &
           SaveCounter, time, nsx_n_cell0, nsx_n_cell, nsx_n_vert,       &
           nsx_nz, No_vp, xvt, yvt, zvt, ufnp, vfnp, wfnp, pfnp) /= 0) then
        write(*,'(a)') ' nsx: SavetecCenter: cannot write ../output/Serial'
        stop 'nsx: fatal error'
      endif
      END SUBROUTINE SavetecCenter

!  SaveParaviewVertex.F90:185-278 -- V-NNNN.vtk of the vertex fields
      SUBROUTINE SaveParaviewVertex(alphafv,ufv,vfv,wfv,pfv,             &
                                    alphasv,usv,vsv,wsv,psv,rhosv,       &
                                    xvt,yvt,zvt,No_vp)
      use nsx_dropin
      implicit none
      real*8, dimension(:,:) :: alphafv, ufv, vfv, wfv, pfv
      real*8, dimension(:,:) :: alphasv, usv, vsv, wsv, psv, rhosv
      real*8, dimension(:,:) :: xvt, yvt, zvt
      integer, dimension(:,:) :: No_vp
      real*8 :: tFin, tIni, dt, tInisave, dtsave, LastTimeSave, time
      integer :: tlimit, SaveCounter, Numictec, nc_ref, FinRK
      common/times/tFin,tIni,dt,tInisave,dtsave,LastTimeSave,time,tlimit
      common/integ/SaveCounter,Numictec,nc_ref,FinRK
      if (nsx_write_vtk_vertex('../output/Serial'//c_null_char,          &
           SaveCounter, nsx_n_cell0, nsx_n_vert, nsx_nz, No_vp,          &
           xvt, yvt, zvt, ufv, vfv, wfv, usv, vsv, wsv) /= 0) then
        write(*,'(a)') ' nsx: SaveParaviewVertex: cannot write ../output/Serial'
        stop 'nsx: fatal error'
      endif
      END SUBROUTINE SaveParaviewVertex

!  SavetecVertexCenter.F90:212-305 -- VC-NNNNN.tec, vertices and centres
      SUBROUTINE SavetecVC(alphafnp,ufnp,vfnp,wfnp,pfnp,                 &
                           alphasnp,usnp,vsnp,wsnp,psnp,rhos,            &
                           xct,yct,zct,No_cp,                            &
                           alphafv,ufv,vfv,wfv,pfv,                      &
                           alphasv,usv,vsv,wsv,psv,rhosv,                &
                           xvt,yvt,zvt,No_vp)
      use nsx_dropin
      implicit none
      real*8, dimension(:,:) :: alphafnp, ufnp, vfnp, wfnp, pfnp
      real*8, dimension(:,:) :: alphasnp, usnp, vsnp, wsnp, psnp, rhos
      real*8, dimension(:,:) :: xct, yct, zct
      integer, dimension(:,:) :: No_cp
      real*8, dimension(:,:) :: alphafv, ufv, vfv, wfv, pfv
      real*8, dimension(:,:) :: alphasv, usv, vsv, wsv, psv, rhosv
      real*8, dimension(:,:) :: xvt, yvt, zvt
      integer, dimension(:,:) :: No_vp
      real*8 :: tFin, tIni, dt, tInisave, dtsave, LastTimeSave, time
      integer :: tlimit, SaveCounter, Numictec, nc_ref, FinRK
      common/times/tFin,tIni,dt,tInisave,dtsave,LastTimeSave,time,tlimit
      common/integ/SaveCounter,Numictec,nc_ref,FinRK
      if (nsx_write_tec_vc('../output/Serial'//c_null_char, SaveCounter, &
           time, nsx_n_cell0, nsx_n_cell, nsx_n_vert, nsx_nz, No_vp,     &
           xvt, yvt, zvt, xct, yct, zct, ufv, vfv, wfv, pfv,             &
           ufnp, vfnp, wfnp, pfnp) /= 0) then
        write(*,'(a)') ' nsx: SavetecVC: cannot write ../output/Serial'
        stop 'nsx: fatal error'
      endif
      END SUBROUTINE SavetecVC
