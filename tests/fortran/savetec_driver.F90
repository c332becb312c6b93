!  savetec_driver.F90 -- test driver of the Fortran snapshot drop-ins
!  (SavetecVertex, SavetecCenter, SaveParaviewVertex, SavetecVC of
!  nsx_dropin.F90), which write files on the
!  host and need no device.
!
!  Reads a stream file written by tests/test_fortran.py: the sizes N_CELL0,
!  N_CELL, N_VERT, NZ, SaveCounter, No_vp (N_CELL0,3), time, xvt, yvt, zvt,
!  ufv, vfv, wfv, pfv (N_VERT,NZ-1), ufnp, vfnp, wfnp, pfnp (N_CELL,NZ) and
!  xct, yct, zct (N_CELL,NZ); the solid-phase vertex velocities are zero.
!  Sets the sizes of the drop-in (nsx_dropin_open; whether a device answers
!  does not matter here), puts SaveCounter and time into the reference's
!  COMMON blocks and calls the four subroutines under the reference's
!  argument lists; the files land in ../output/Serial.
      program savetec_driver
      use iso_c_binding
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
        SUBROUTINE SaveParaviewVertex(alphafv,ufv,vfv,wfv,pfv,           &
                                      alphasv,usv,vsv,wsv,psv,rhosv,     &
                                      xvt,yvt,zvt,No_vp)
          real*8, dimension(:,:) :: alphafv, ufv, vfv, wfv, pfv
          real*8, dimension(:,:) :: alphasv, usv, vsv, wsv, psv, rhosv
          real*8, dimension(:,:) :: xvt, yvt, zvt
          integer, dimension(:,:) :: No_vp
        END SUBROUTINE SaveParaviewVertex
        SUBROUTINE SavetecVC(alphafnp,ufnp,vfnp,wfnp,pfnp,               &
                             alphasnp,usnp,vsnp,wsnp,psnp,rhos,          &
                             xct,yct,zct,No_cp,                          &
                             alphafv,ufv,vfv,wfv,pfv,                    &
                             alphasv,usv,vsv,wsv,psv,rhosv,              &
                             xvt,yvt,zvt,No_vp)
          real*8, dimension(:,:) :: alphafnp, ufnp, vfnp, wfnp, pfnp
          real*8, dimension(:,:) :: alphasnp, usnp, vsnp, wsnp, psnp, rhos
          real*8, dimension(:,:) :: xct, yct, zct
          integer, dimension(:,:) :: No_cp
          real*8, dimension(:,:) :: alphafv, ufv, vfv, wfv, pfv
          real*8, dimension(:,:) :: alphasv, usv, vsv, wsv, psv, rhosv
          real*8, dimension(:,:) :: xvt, yvt, zvt
          integer, dimension(:,:) :: No_vp
        END SUBROUTINE SavetecVC
      end interface
!     COMMON /times/ and /integ/ of the reference (common.mpf:221-228,310-346)
      real*8 :: tFin, tIni, dt, tInisave, dtsave, LastTimeSave, time
      integer :: tlimit, SaveCounter, Numictec, nc_ref, FinRK
      common/times/tFin,tIni,dt,tInisave,dtsave,LastTimeSave,time,tlimit
      common/integ/SaveCounter,Numictec,nc_ref,FinRK
      character(len=1024) :: fin
      integer :: u, n0, nc, nv, nz, rc
      integer, allocatable :: No_vp(:,:)
      real*8, allocatable :: xvt(:,:), yvt(:,:), zvt(:,:), fv(:,:,:), fc(:,:,:)
      real*8, allocatable :: xct(:,:), yct(:,:), zct(:,:), sv(:,:)
      integer, allocatable :: No_cp(:,:)

      call get_command_argument(1, fin)
      open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
      read(u) n0, nc, nv, nz, SaveCounter
      allocate(No_vp(n0,3), xvt(nv,nz-1), yvt(nv,nz-1), zvt(nv,nz-1))
      allocate(fv(nv,nz-1,4), fc(nc,nz,4))
      allocate(xct(nc,nz), yct(nc,nz), zct(nc,nz), sv(nv,nz-1), No_cp(nc,3))
      read(u) No_vp, time, xvt, yvt, zvt, fv, fc, xct, yct, zct
      close(u)
      sv = 0.d0
      No_cp = 0
      tFin = -1.d0
      tlimit = -1
      rc = nsx_dropin_open(n0, nc, nv, nz)
      call SavetecVertex(fv(:,:,1), fv(:,:,1), fv(:,:,2), fv(:,:,3),      &
                         fv(:,:,4), fv(:,:,1), fv(:,:,1), fv(:,:,1),      &
                         fv(:,:,1), fv(:,:,1), fv(:,:,1), xvt, yvt, zvt,  &
                         No_vp)
      call SavetecCenter(fc(:,:,1), fc(:,:,1), fc(:,:,2), fc(:,:,3),      &
                         fc(:,:,4), fc(:,:,1), fc(:,:,1), fc(:,:,1),      &
                         fc(:,:,1), fc(:,:,1), fc(:,:,1), xvt, yvt, zvt,  &
                         No_vp)
      call SaveParaviewVertex(fv(:,:,1), fv(:,:,1), fv(:,:,2), fv(:,:,3), &
                              fv(:,:,4), sv, sv, sv, sv, sv, sv,          &
                              xvt, yvt, zvt, No_vp)
      call SavetecVC(fc(:,:,1), fc(:,:,1), fc(:,:,2), fc(:,:,3),          &
                     fc(:,:,4), fc(:,:,1), fc(:,:,1), fc(:,:,1),          &
                     fc(:,:,1), fc(:,:,1), fc(:,:,1), xct, yct, zct,      &
                     No_cp, fv(:,:,1), fv(:,:,1), fv(:,:,2), fv(:,:,3),   &
                     fv(:,:,4), sv, sv, sv, sv, sv, sv, xvt, yvt, zvt,    &
                     No_vp)
      write(*,'(a,i4)') ' nsx savetec driver: snapshots written, counter', SaveCounter
      if (rc == 0) call nsx_dropin_close()
      end program savetec_driver
