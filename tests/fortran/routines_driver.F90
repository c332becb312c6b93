!  routines_driver.F90 -- test driver of the single-routine drop-ins of
!  nsx_dropin.F90 (the routines of src/AdvectionDiffusion, src/Gradient,
!  src/LES, src/Interpolation, src/BoundaryConditions, BICGSTAB_SOLVE and
!  src/LoopUpdate), called under the reference's names and argument lists.
!
!  Reads the case file of dropin_driver.F90 (tests/test_fortran.py writes
!  it), sets the device up and, on the initial fields u, v, w, p:
!    BCglobalVC(u,v,w -> uv,vv,wv, tagdir 2); BCglobalP(p -> pv, tagdir 3);
!    gradientP(p -> dpdx,dpdy,dpds); calcul_face(u,uv -> 5 faces);
!    interpolateU(u,v,w -> 5 faces); sgles(u,v,w) -> eddy;
!    Source3D(u,v,w -> src, resident faces);
!    AdvDiffNS(bdx,bdy,bdz += ..., Gam = eddy + 1/Re);
!    update_rhsp(src/dt, p -> rhs, pv2); newu3D(p -> resident faces, pv3);
!    BICGSTAB_SOLVE(AmC*, rhs mean removed, phi = 0);
!    LoopTimeInitial and LoopTimeUpdate on level arrays.
!  Writes every result, in that order, to the output stream file.  Exit
!  status 3 when no HIP device is usable.
      program routines_driver
      use iso_c_binding
      use nsx_iface
      use nsx_dropin
      implicit none
      interface
        SUBROUTINE LoopTimeInitial(etanp,alphafnp,ufnp,vfnp,wfnp,pfnp,   &
             alphasnp,usnp,vsnp,wsnp,psnp,xct,yct,zct,xvt,yvt,zvt,etan,  &
             alphafn,ufn,vfn,wfn,pfn,alphasn,usn,vsn,wsn,psn,xc,yc,sig,   &
             Hpr,h,xv,yv,sigv,Hprv,hv,No_cp,nbe)
          real*8, dimension(:) :: etanp, etan
          real*8, dimension(:,:) :: alphafnp,ufnp,vfnp,wfnp,pfnp,alphasnp,usnp,vsnp,wsnp,psnp
          real*8, dimension(:,:) :: xct,yct,zct,xvt,yvt,zvt
          real*8, dimension(:,:) :: alphafn,ufn,vfn,wfn,pfn,alphasn,usn,vsn,wsn,psn
          real*8, dimension(:) :: xc,yc,sig,Hpr,h,xv,yv,sigv,Hprv,hv
          integer, dimension(:,:) :: No_cp
          integer, dimension(:) :: nbe
        END SUBROUTINE LoopTimeInitial
        SUBROUTINE LoopTimeUpdate(etan,alphafn,ufn,vfn,wfn,pfn,          &
             alphasn,usn,vsn,wsn,psn,xct,yct,zct,xvt,yvt,zvt,etanp,      &
             alphafnp,ufnp,vfnp,wfnp,pfnp,alphasnp,usnp,vsnp,wsnp,psnp,  &
             xc,yc,sig,Hpr,h,xv,yv,sigv,Hprv,hv,No_cp,nbe)
          real*8, dimension(:) :: etanp, etan
          real*8, dimension(:,:) :: alphafnp,ufnp,vfnp,wfnp,pfnp,alphasnp,usnp,vsnp,wsnp,psnp
          real*8, dimension(:,:) :: xct,yct,zct,xvt,yvt,zvt
          real*8, dimension(:,:) :: alphafn,ufn,vfn,wfn,pfn,alphasn,usn,vsn,wsn,psn
          real*8, dimension(:) :: xc,yc,sig,Hpr,h,xv,yv,sigv,Hprv,hv
          integer, dimension(:,:) :: No_cp
          integer, dimension(:) :: nbe
        END SUBROUTINE LoopTimeUpdate
      end interface
      character(len=1024) :: fin, fout
      integer :: u, n0, nc, nv, nz, rc, i, k
      integer(c_int), allocatable :: No_cp(:,:), No_vp(:,:), nbe(:), nbev(:)
      integer(c_int), allocatable :: surrounding(:,:), Dimsurrounding(:), vertex_pair(:), vertex_edge(:)
      integer(c_int), allocatable :: bcc(:,:), bcv(:,:)
      integer(c_int) :: ipar(7)
      real(c_double), allocatable :: areaCell(:), e3(:,:,:), lsm(:,:), sig(:), dsig(:), sigv(:), dsigv(:)
      real(c_double), allocatable :: weight(:,:), dlVsum(:), Am(:,:,:), walld(:)
      real(c_double) :: rpar(9), avec
      real(c_double), allocatable :: f0(:,:,:), rhs(:,:), xvt(:,:), yvt(:,:), zvt(:,:)
      real(c_double), allocatable :: vtx(:,:,:), grd(:,:,:), fc(:,:,:), fu(:,:,:), eddy(:,:), src(:,:)
      real(c_double), allocatable :: rb(:,:,:), gam(:,:), r2(:,:), pv2(:,:), pv3(:,:), nf(:,:,:), phi(:,:)
      real(c_double), allocatable :: xc(:), xv(:), sv(:), one(:), onev(:), zc(:), zv(:)
      real(c_double), allocatable :: lev(:,:,:), lnp(:,:,:), eta(:), etanp(:), cct(:,:,:), cvt(:,:,:)

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
      read(u) n0, nc, nv, nz
      allocate(No_cp(nc,3), No_vp(n0,3), nbe(n0), nbev(nv), surrounding(nv,10), Dimsurrounding(nv))
      allocate(vertex_pair(nv), vertex_edge(nv), bcc(nc,4), bcv(nv,4))
      read(u) No_cp, No_vp, nbe, nbev, surrounding, Dimsurrounding, vertex_pair, vertex_edge, bcc, bcv, ipar
      allocate(areaCell(nc), e3(n0,3,9), lsm(n0,3), sig(nz), dsig(nz), sigv(nz-1), dsigv(nz-1))
      allocate(weight(nv,10), dlVsum(nv), Am(n0,nz,6), walld(nz))
      read(u) areaCell, e3, lsm, sig, dsig, sigv, dsigv, weight, dlVsum, Am, walld, rpar
      allocate(f0(nc,nz,4), rhs(nc,nz))
      read(u) f0, rhs
      allocate(xvt(nv,nz-1), yvt(nv,nz-1), zvt(nv,nz-1))
      read(u) xvt, yvt, zvt
      close(u)

      rc = nsx_dropin_open(n0, nc, nv, nz)
      if (rc /= 0) then
        write(*,'(a)') ' nsx: no usable HIP device'
        call exit(3)
      endif
      call nsx_dropin_setup(No_cp, No_vp, nbe, nbev,                     &
           areaCell, e3(:,:,1), e3(:,:,2), e3(:,:,3), e3(:,:,4),          &
           e3(:,:,5), e3(:,:,6), e3(:,:,7), e3(:,:,8), e3(:,:,9),         &
           lsm(:,1), lsm(:,2), lsm(:,3), rpar(1),                         &
           sig, dsig, sigv, dsigv, rpar(2), rpar(3),                      &
           surrounding, weight, Dimsurrounding, dlVsum, vertex_pair,      &
           vertex_edge, bcc(:,1), bcc(:,2), bcc(:,3), bcc(:,4),           &
           bcv(:,1), bcv(:,2), bcv(:,3), bcv(:,4),                        &
           ipar(1), ipar(2), ipar(3), ipar(4),                            &
           Am(:,:,1), Am(:,:,2), Am(:,:,3), Am(:,:,4), Am(:,:,5),         &
           Am(:,:,6), walld, rpar(4), rpar(5), rpar(6), rpar(7), rpar(8), &
           rpar(9), ipar(5), ipar(6), ipar(7))
!     the geometry arguments the device path does not read again
      allocate(xc(nc), xv(nv), sv(nz-1), one(nc), onev(nv), zc(nc), zv(nv))
      xc = 0.d0; xv = 0.d0; sv = 0.d0; one = 1.d0; onev = 1.d0; zc = 0.d0; zv = 0.d0
      allocate(vtx(nv,nz-1,4), grd(nc,nz,3), fc(n0,nz,5), fu(n0,nz,5), eddy(nc,nz), src(nc,nz))
      allocate(rb(nc,nz,3), gam(nc,nz), r2(nc,nz), pv2(nv,nz-1), pv3(nv,nz-1), nf(n0,nz,5), phi(nc,nz))
      vtx = 0.d0; grd = 0.d0; fc = 0.d0; fu = 0.d0; src = 0.d0; r2 = 0.d0; pv2 = 0.d0; pv3 = 0.d0

      call BCglobalVC(f0(:,:,1), f0(:,:,2), f0(:,:,3),                  &
                      vtx(:,:,1), vtx(:,:,2), vtx(:,:,3),                &
                      xc, xc, sig, dsig, No_cp, nbe,                     &
                      xv, xv, sigv, dsigv, No_vp, nbev, 2)
      call BCglobalP(f0(:,:,4), vtx(:,:,4), xc, xc, sig, dsig, No_cp,    &
                     nbe, xv, xv, sigv, dsigv, No_vp, nbev, 3)
      call gradientP(f0(:,:,4), vtx(:,:,4), grd(:,:,1), grd(:,:,2),      &
                     grd(:,:,3), xc, xc, sig, dsig, No_cp, nbe,          &
                     xv, xv, sigv, dsigv, No_vp, nbev)
      call calcul_face(fc(:,:,1), fc(:,:,2), fc(:,:,3), fc(:,:,4),       &
                       fc(:,:,5), f0(:,:,1), vtx(:,:,1),                 &
                       xc, xc, sig, dsig, No_cp, nbe,                    &
                       xv, xv, sigv, dsigv, No_vp, nbev, 1)
      call interpolateU(fu(:,:,1), fu(:,:,2), fu(:,:,3), fu(:,:,4),      &
                        fu(:,:,5), f0(:,:,1), f0(:,:,2), f0(:,:,3),      &
                        xc, xc, sig, dsig, No_cp, nbe,                   &
                        xv, xv, sigv, dsigv, No_vp, nbev)
      call sgles(f0(:,:,1), f0(:,:,2), f0(:,:,3),                        &
                 xc, xc, sig, dsig, No_cp, nbe,                          &
                 xv, xv, sigv, dsigv, No_vp, nbev,                       &
                 one, zc, zc, zc, onev, zv, zv)
      call nsx_dropin_get(NSX_F_EDDY, eddy)
      call Source3D(src, f0(:,:,1), f0(:,:,2), f0(:,:,3),                &
                    xc, xc, sig, dsig, No_cp, nbe,                       &
                    xv, xv, sigv, dsigv, No_vp, nbev)
!     hydro.F90:280-293: Gam = eddy + 1/Re; the rates start from the forcing
      do k = 1, nz
        do i = 1, nc
          gam(i,k) = eddy(i,k)+1./rpar(5)
        enddo
      enddo
      rb(:,:,1) = rpar(6)
      rb(:,:,2) = rpar(7)
      rb(:,:,3) = rpar(8)
      call AdvDiffNS(rb(:,:,1), rb(:,:,2), rb(:,:,3),                    &
                     f0(:,:,1), f0(:,:,2), f0(:,:,3), gam, gam, gam,     &
                     xc, xc, sig, dsig, No_cp, nbe,                      &
                     xv, xv, sigv, dsigv, No_vp, nbev)
      r2 = src / rpar(4)
      call update_rhsp(r2, f0(:,:,4), pv2, grd(:,:,1), grd(:,:,2),       &
                       grd(:,:,3), xc, xc, sig, dsig, No_cp, nbe,        &
                       xv, xv, sigv, dsigv, No_vp, nbev)
      call newu3D(f0(:,:,4), pv3, 0, grd(:,:,1), grd(:,:,2), grd(:,:,3), &
                  xc, xc, sig, dsig, No_cp, nbe,                         &
                  xv, xv, sigv, dsigv, No_vp, nbev)
      call nsx_dropin_get(NSX_F_U1FACE, nf(:,:,1))
      call nsx_dropin_get(NSX_F_U2FACE, nf(:,:,2))
      call nsx_dropin_get(NSX_F_U3FACE, nf(:,:,3))
      call nsx_dropin_get(NSX_F_UTFACE, nf(:,:,4))
      call nsx_dropin_get(NSX_F_UBFACE, nf(:,:,5))
!     BICGSTAB_SOLVE on the mean-removed right-hand side (Glob_Norm(rhs,1),
!     bicgstab.F90:635-663), from phi = 0
      avec = 0.d0
      do k = 2, nz-1
        do i = 1, n0
          avec = avec + rhs(i,k)
        enddo
      enddo
      avec = avec/(rpar(1)*(rpar(3)-rpar(2)))
      do k = 2, nz-1
        do i = 1, n0
          rhs(i,k) = rhs(i,k) - avec*areaCell(i)*dsig(k)
        enddo
      enddo
      phi = 0.d0
      call BICGSTAB_SOLVE(Am(:,:,1), Am(:,:,2), Am(:,:,3), Am(:,:,4),    &
                          Am(:,:,5), Am(:,:,6), rhs, phi,                &
                          xc, xc, sig, dsig, No_cp, nbe, 1)
!     LoopTimeInitial then LoopTimeUpdate on five level fields
      allocate(lev(nc,nz,10), lnp(nc,nz,10), eta(nc), etanp(nc), cct(nc,nz,3), cvt(nv,nz-1,3))
      do k = 1, 10
        lev(:,:,k) = dble(k)
        lnp(:,:,k) = -dble(k)
      enddo
      eta = 0.5d0
      etanp = -0.5d0
      call LoopTimeInitial(etanp, lnp(:,:,1), lnp(:,:,2), lnp(:,:,3),    &
           lnp(:,:,4), lnp(:,:,5), lnp(:,:,6), lnp(:,:,7), lnp(:,:,8),    &
           lnp(:,:,9), lnp(:,:,10), cct(:,:,1), cct(:,:,2), cct(:,:,3),   &
           cvt(:,:,1), cvt(:,:,2), cvt(:,:,3), eta, lev(:,:,1),           &
           lev(:,:,2), lev(:,:,3), lev(:,:,4), lev(:,:,5), lev(:,:,6),    &
           lev(:,:,7), lev(:,:,8), lev(:,:,9), lev(:,:,10), xc, xc, sig,  &
           one, zc, xv, xv, sigv, onev, zv, No_cp, nbe)
      lnp = lnp + 1.d0
      call LoopTimeUpdate(eta, lev(:,:,1), lev(:,:,2), lev(:,:,3),       &
           lev(:,:,4), lev(:,:,5), lev(:,:,6), lev(:,:,7), lev(:,:,8),    &
           lev(:,:,9), lev(:,:,10), cct(:,:,1), cct(:,:,2), cct(:,:,3),   &
           cvt(:,:,1), cvt(:,:,2), cvt(:,:,3), etanp, lnp(:,:,1),         &
           lnp(:,:,2), lnp(:,:,3), lnp(:,:,4), lnp(:,:,5), lnp(:,:,6),    &
           lnp(:,:,7), lnp(:,:,8), lnp(:,:,9), lnp(:,:,10), xc, xc, sig,  &
           one, zc, xv, xv, sigv, onev, zv, No_cp, nbe)

      open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(u) f0, vtx, grd, fc, fu, eddy, src, rb, r2, pv2, pv3, nf, phi, lev, eta, cct, cvt
      close(u)
      write(*,'(a,i6,a)') ' nsx routines drop-in: done (', nsx_last_info%iters, ' iterations in BICGSTAB_SOLVE)'
      call nsx_dropin_close()
      end program routines_driver
