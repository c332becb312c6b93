!  dropin_driver.F90 -- test driver of the Fortran drop-in (nsx_dropin.F90).
!
!  Reads a case written by tests/test_fortran.py (stream file: sizes, the
!  reference's set-up arrays, the initial fields and a PPE right-hand side),
!  sets the device up through nsx_dropin_setup and calls the drop-in
!  subroutines under the reference's names and argument lists: bicgstab on the
!  given right-hand side, then one hydro stage (flag_ab = 0) from the initial
!  fields; then the channel statistics: sm_init, glob_sample of the initial
!  fields and of the stage result, glob_put_plane_mm (its files go to
!  ../output/Stats of the working directory); then the Tecplot snapshots
!  SavetecVertex of the stage's vertex fields and SavetecCenter of its cell
!  fields, with SaveCounter = 7 and time = 0.0125 in the reference's COMMON
!  blocks (files in ../output/Serial).  Writes phi, phiv, ufnp, vfnp, wfnp,
!  pfnp, ufv, vfv, wfv, pfv to the output stream file.  Exit status 3 when no
!  HIP device is usable.
      program dropin_driver
      use iso_c_binding
      use nsx_iface
      use nsx_dropin
      implicit none
!     the explicit interfaces of the snapshot writers (module_interfaces.F90:141-190)
      interface
        SUBROUTINE SavetecVertex(alphafv,ufv,vfv,wfv,pfv,                &
                                 alphasv,usv,vsv,wsv,psv,rhosv,          &
                                 xvt,yvt,zvt,No_vp)
          real*8, dimension(:,:) :: alphafv, ufv, vfv, wfv, pfv
          real*8, dimension(:,:) :: alphasv, usv, vsv, wsv, psv, rhosv
          real*8, dimension(:,:) :: xvt, yvt, zvt
          integer, dimension(:,:) :: No_vp
        END SUBROUTINE SavetecVertex
        SUBROUTINE SavetecCenter(alphafnp,ufnp,vfnp,wfnp,pfnp,           &
                                 alphasnp,usnp,vsnp,wsnp,psnp,rhos,      &
                                 xvt,yvt,zvt,No_vp)
          real*8, dimension(:,:) :: alphafnp, ufnp, vfnp, wfnp, pfnp
          real*8, dimension(:,:) :: alphasnp, usnp, vsnp, wsnp, psnp, rhos
          real*8, dimension(:,:) :: xvt, yvt, zvt
          integer, dimension(:,:) :: No_vp
        END SUBROUTINE SavetecCenter
      end interface
!     COMMON /times/ and /integ/ of the reference (common.mpf:221-228,310-346)
      real*8 :: tFin, tIni, dt, tInisave, dtsave, LastTimeSave, time
      integer :: tlimit, SaveCounter, Numictec, nc_ref, FinRK
      common/times/tFin,tIni,dt,tInisave,dtsave,LastTimeSave,time,tlimit
      common/integ/SaveCounter,Numictec,nc_ref,FinRK
      real(c_double), allocatable :: xvt(:,:), yvt(:,:), zvt(:,:)
      character(len=1024) :: fin, fout
      integer :: u, n0, nc, nv, nz, rc
      integer(c_int), allocatable :: No_cp(:,:), No_vp(:,:), nbe(:), nbev(:)
      integer(c_int), allocatable :: surrounding(:,:), Dimsurrounding(:), vertex_pair(:), vertex_edge(:)
      integer(c_int), allocatable :: bcc(:,:), bcv(:,:)
      integer(c_int) :: ipar(7)
      real(c_double), allocatable :: areaCell(:), e3(:,:,:), lsm(:,:), sig(:), dsig(:), sigv(:), dsigv(:)
      real(c_double), allocatable :: weight(:,:), dlVsum(:), Am(:,:,:), walld(:)
      real(c_double) :: rpar(9)
      real(c_double), allocatable :: f0(:,:,:), rhs(:,:), phi(:,:), phiv(:,:)
      real(c_double), allocatable :: fnp(:,:,:), fv(:,:,:), rho(:,:), rhov(:,:), xc(:), xv(:), sv(:)
      integer, allocatable :: nbei(:)

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
!     ipar: bctop, bcbot, TopBC, BotBC, MaxIters, FinRK, Iturbu
!     rpar: TotalVolume, sigIni, sigFin, dt, Re, bdx, bdy, bdz, eps
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

!     the arguments the reference passes but the device path does not read
      allocate(xc(nc), xv(nv), sv(nz-1), nbei(n0), rho(nc,nz), rhov(nv,nz-1))
      xc = 0.d0; xv = 0.d0; sv = 0.d0; nbei = 0; rho = 1.d0; rhov = 1.d0

!     bicgstab(phi,phiv,rhs, ...) from phi = 0
      allocate(phi(nc,nz), phiv(nv,nz-1))
      phi = 0.d0
      call bicgstab(phi, phiv, rhs, xc, xc, sig, dsig, No_cp, nbe,       &
                    xv, xv, sigv, dsigv, No_vp, nbev, xc, xc, xc, xv, xv, xv)

!     the initial state, then one hydro stage from it (stage input = level n)
      call nsx_dropin_put(NSX_F_UFN, f0(:,:,1))
      call nsx_dropin_put(NSX_F_VFN, f0(:,:,2))
      call nsx_dropin_put(NSX_F_WFN, f0(:,:,3))
      call nsx_dropin_put(NSX_F_PFN, f0(:,:,4))
      call nsx_dropin_check(nsx_init_state(nsx_ctx), 'nsx_init_state')
      allocate(fnp(nc,nz,4), fv(nv,nz-1,4))
      call hydro(fnp(:,:,1), f0(:,:,1), f0(:,:,1), fv(:,:,1),            &
                 fnp(:,:,2), f0(:,:,2), f0(:,:,2), fv(:,:,2),            &
                 fnp(:,:,3), f0(:,:,3), f0(:,:,3), fv(:,:,3),            &
                 fnp(:,:,4), f0(:,:,4), f0(:,:,4), fv(:,:,4),            &
                 rho, rho, rhov, rhov,                                   &
                 xc, xc, sig, dsig, No_cp, nbe,                          &
                 xv, xv, sigv, dsigv, No_vp, nbev,                       &
                 xc, xc, xc, xc, xv, xv, xv, 0)

      call sm_init()
      call glob_sample(f0(:,:,1), f0(:,:,2), f0(:,:,3), f0(:,:,4))
      call glob_sample(fnp(:,:,1), fnp(:,:,2), fnp(:,:,3), fnp(:,:,4))
      call glob_put_plane_mm()

!     the snapshots of the stage (SaveCounter and time from the COMMON blocks)
      SaveCounter = 7
      time = 0.0125d0
      tFin = -1.d0
      tlimit = -1
      call SavetecVertex(rhov, fv(:,:,1), fv(:,:,2), fv(:,:,3), fv(:,:,4),&
                         rhov, rhov, rhov, rhov, rhov, rhov,              &
                         xvt, yvt, zvt, No_vp)
      call SavetecCenter(rho, fnp(:,:,1), fnp(:,:,2), fnp(:,:,3),          &
                         fnp(:,:,4), rho, rho, rho, rho, rho, rho,         &
                         xvt, yvt, zvt, No_vp)

      open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(u) phi, phiv, fnp, fv
      close(u)
      write(*,'(a,i6,a)') ' nsx drop-in: bicgstab and hydro done (', nsx_last_info%iters, ' iterations in the stage solve)'
      call nsx_dropin_close()
      end program dropin_driver
