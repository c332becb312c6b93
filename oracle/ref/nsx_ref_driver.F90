!  nsx_ref_driver -- runs the reference's own sequential routines on a case
!  written by the nsx set-up (tests/golden and oracle tooling), and dumps
!  what they produce.
!
!  It replaces only the file parsing and the random initial condition of the
!  reference PROGRAM (src/nsmp3D.F90:116-304): the mesh, parameters, sigma
!  levels and initial field are read from a binary case file; everything else
!  is the reference's own code, called in the order of src/nsmp3D.F90 and
!  src/Main/user_initial.F90.  Test infrastructure: compiled and run only in
!  the build container.
!
!  usage: nsx_ref_driver <case.bin> <out.bin>
!  mode 3 dumps the set-up geometry alone (with hdr(11) = 1 the mesh comes
!  from data.txt through the reference's input.F90);
!  mode 1 dumps set-up geometry and single-routine results on the initial
!  field; mode 2 runs nsteps RK steps, dumps the state after each step and
!  calls glob_cfl on the final state (its printed diagnostics and meanuv.dat
!  are the record of that call).  Mode 2 also takes the channel statistics
!  as nsmp3D.F90:303,579-589 does (sm_init, then glob_sample after every
!  step, glob_put_plane_mm at the end, whose files go to ../output/Stats of
!  the working directory) and dumps the twelve accumulators; then it saves
!  the final state as nsmp3D.F90:617-625 does for ChooseExit 1 and 2
!  (SavetecVertex, SavetecCenter, SaveParaviewVertex, SavetecVC, SaveCounter
!  3, into ../output/Serial) and dumps the coordinates xvt, yvt, zvt, xct,
!  yct, zct and the solid-phase vertex velocities those files hold.

      module nsxdump
      implicit none
      integer :: udump
      contains
      subroutine dump_r8(name, a, n)
      character(*) :: name
      integer :: n
      real*8 :: a(n)
      character(len=32) :: nm
      nm = name
      write(udump) nm, int(1,4), int(n,8), a
      end subroutine
      subroutine dump_i4(name, a, n)
      character(*) :: name
      integer :: n
      integer :: a(n)
      character(len=32) :: nm
      nm = name
      write(udump) nm, int(2,4), int(n,8), a
      end subroutine
      end module nsxdump

      program nsx_ref_driver
      USE variables
      USE geometry
      USE les
      USE stats
      USE nsxdump
!     explicit interfaces of the routines with assumed-shape dummies, as the
!     reference PROGRAM has them (nsmp3D.F90:55); grandientLSM's entry there
!     does not match the routine and is left out
      USE interfaces, only: input, calcul_geometry, initial, LoopTimeInitial, LoopTimeUpdate, &
                            SavetecVertex, SavetecCenter, SaveParaviewVertex, SavetecVC
      implicit none
      character(len=512) :: casefile, outfile
      integer :: hdr(16), ucase, n1, n2, nsteps, mode, irk, istep, nbuf
      real*8 :: rhdr(16)
      real*8, allocatable :: buf(:)
      real*8, allocatable :: w1(:,:), w2(:,:), w3(:,:), w4(:,:), w5(:,:), w6(:,:)
      real*8, allocatable :: g1(:,:), g2(:,:), g3(:,:), t1(:,:), t2(:,:), t3(:,:)
      real*8, allocatable :: v1(:,:), v2(:,:), v3(:,:), v4(:,:)
      real*8, allocatable :: f1(:,:), f2(:,:), f3(:,:), f4(:,:), f5(:,:)
      real*8 :: tv(1)
      integer :: iv(4)
      integer(8) :: clk0, clk1, clkrate

      call get_command_argument(1, casefile)
      call get_command_argument(2, outfile)
      open(newunit=ucase, file=trim(casefile), access='stream', form='unformatted', status='old')
      read(ucase) hdr, rhdr
      if (hdr(1) .ne. 1314084913) stop 'bad case file'
      if (hdr(4) .ne. NN) then
         print*, 'case NN', hdr(4), ' but this build has NN', NN
         stop 2
      endif
      mode   = hdr(9)
      nsteps = hdr(5)

!     ---- sizes (nsmp3D.F90:128-133)
      NZ      = NZglobal
      N_VERT  = hdr(3)
      N_CELL0 = hdr(2)
      N_VERTglobal  = N_VERT
      N_CELL0global = N_CELL0
      N_CELL  = N_CELL0 + N_CELLghostMAX

!     ---- parameters normally read by input_parameters.F90
      FinRK = hdr(6)
      MaxIters = hdr(7)
      Iturbu = hdr(8)
      ChooseExit = 0
      Ipress = 1
      Icase = 0
      Ihydro = 0
      Idepth = 1
      Iinit0 = 0
      tIni = 0.0d0
      dt = rhdr(1)
      Re = rhdr(2)
      bdx = rhdr(3)
      bdy = rhdr(4)
      bdz = rhdr(5)
      eps = rhdr(6)
      tFin = dble(nsteps)*dt
      ncfl = 1000000
      tlimit = 1000000
      h0 = 0.0d0

      call alloc_variables
      call alloc_geometry

!     ---- mesh (what input.F90 reads from data.txt)
      call read_i4(ucase, No_vp, 3*N_CELL0)
      call read_cp(ucase)
      call read_i4(ucase, nbe, N_CELL0)
      call read_i4(ucase, nbev, N_VERT)
      read(ucase) (xv(n1), n1=1,N_VERT)
      read(ucase) (yv(n1), n1=1,N_VERT)
      read(ucase) (zbv(n1), n1=1,N_VERT)
!     hdr(11) = 1: the mesh from data.txt in the working directory instead,
!     read by the reference's own input.F90
      if (hdr(11) .eq. 1) call input(No_vp,No_cp,nbe,nbev,xv,yv,zbv)
      call system_clock(clk0, clkrate)
      do n1=1,N_VERT
         hv(n1) = 0.0d0
      enddo
      do n1=1,N_CELL0
         h(n1) = 0.0d0
      enddo

!     ---- geometry, BC flags, PPE matrix (nsmp3D.F90:232-249)
      call calcul_geometry(xc,yc,sig,dsig,No_cp,nbe,h, &
                           xv,yv,sigv,dsigv,No_vp,nbev,&
                           ic1tec,ic2tec,ic3tec)
!     sigma levels of the case replace the uniform ones of calcul_geometry.F90:263-283
      read(ucase) (sig(n1), n1=1,NZ)
      read(ucase) (dsig(n1), n1=1,NZ)
      read(ucase) (sigv(n1), n1=1,NZ-1)
      read(ucase) (dsigv(n1), n1=1,NZ-1)
      call calcul_bc(xc,yc,sig,dsig,No_cp,nbe, &
                     xv,yv,sigv,dsigv,No_vp,nbev)
      call PDE_center(xc,yc,sig,dsig,No_cp,nbe,    &
                      xv,yv,sigv,dsigv,No_vp,nbev)
      call system_clock(clk1)
      write(*,'(a,f12.4)') ' NSX_SETUP_SECONDS', dble(clk1-clk0)/dble(clkrate)

      time  = 0.0d0
      nstep = 0
      call initial(alphafn,ufn,vfn,wfn,pfn,viscof,rhof,   &
                   alphasn,usn,vsn,wsn,psn,viscos,rhos,   &
                   alphafv,ufv,vfv,wfv,pfv,viscofv,rhofv, &
                   alphasv,usv,vsv,wsv,psv,viscosv,rhosv, &
                   etan,etav,Hpr,Hprv,                    &
                   xct,yct,zct,                           &
                   xvt,yvt,zvt,                           &
                   xc,yc,sig,dsig,No_cp,nbe,              &
                   xv,yv,sigv,dsigv,No_vp,nbev,           &
                   h,hv)

!     ---- initial field (user_initial.F90:155-253 with the case's values)
      do n1=1,N_CELL
         h(n1) = 0.0d0
         Hpr(n1) = 1.0d0
      enddo
      do n1=1,N_VERT
         hv(n1) = 0.0d0
         Hprv(n1) = 1.0d0
      enddo
      nbuf = N_CELL0*NZ
      allocate(buf(nbuf))
      read(ucase) buf
      call put_field(ufn, buf)
      read(ucase) buf
      call put_field(vfn, buf)
      read(ucase) buf
      call put_field(wfn, buf)
      read(ucase) buf
      call put_field(pfn, buf)
      close(ucase)
      do n2=1,NZ-1
         do n1=1,N_VERT
            ufv(n1,n2) = 0.0d0
            vfv(n1,n2) = 0.0d0
            wfv(n1,n2) = 0.0d0
            pfv(n1,n2) = 0.0d0
         enddo
      enddo
      call BCglobalVC(ufn,vfn,wfn,              &
                      ufv,vfv,wfv,              &
                      xc,yc,sig,dsig,No_cp,nbe, &
                      xv,yv,sigv,dsigv,No_vp,nbev,3)
      call BCglobalP(pfn,pfv,                  &
                     xc,yc,sig,dsig,No_cp,nbe, &
                     xv,yv,sigv,dsigv,No_vp,nbev,3)
      call interpolateU(U1FACE,U2FACE,U3FACE,UTFACE,UBFACE, &
                        ufn,vfn,wfn,                        &
                        xc,yc,sig,dsig,No_cp,nbe,           &
                        xv,yv,sigv,dsigv,No_vp,nbev)
      allocate(rhsuf(N_CELL,NZ),rhsvf(N_CELL,NZ),rhswf(N_CELL,NZ))
      do n2=1,NZ
         do n1=1,N_CELL
            rhsuf(n1,n2) = 0.0d0
            rhsvf(n1,n2) = 0.0d0
            rhswf(n1,n2) = 0.0d0
         enddo
      enddo

      open(newunit=udump, file=trim(outfile), access='stream', form='unformatted', status='replace')
      iv(1) = N_CELLexact
      iv(2) = N_CELLghost
      iv(3) = bctop
      iv(4) = bcbot
      call dump_i4('sizes', iv, 4)
      tv(1) = TotalVolume
      call dump_r8('TotalVolume', tv, 1)

      if (mode .eq. 1) call unit_mode
      if (mode .eq. 2) call step_mode
      if (mode .eq. 3) call setup_mode
      close(udump)

      contains

      subroutine read_i4(u, a, n)
      integer :: u, n
      integer :: a(n)
      read(u) a
      end subroutine

      subroutine read_cp(u)
      integer :: u, m
      integer, allocatable :: t(:)
      allocate(t(3*N_CELL0))
      read(u) t
      do m=1,N_CELL0
         No_cp(m,1) = t(m)
         No_cp(m,2) = t(N_CELL0+m)
         No_cp(m,3) = t(2*N_CELL0+m)
      enddo
      deallocate(t)
      end subroutine

      subroutine put_field(a, b)
      real*8 :: a(N_CELL,NZ), b(N_CELL0,NZ)
      integer :: m1, m2
      do m2=1,NZ
         do m1=1,N_CELL
            a(m1,m2) = 0.0d0
         enddo
         do m1=1,N_CELL0
            a(m1,m2) = b(m1,m2)
         enddo
      enddo
      end subroutine

      subroutine dump_geometry
      call dump_r8('xc', xc, N_CELL)
      call dump_r8('yc', yc, N_CELL)
      call dump_r8('areaCell', areaCell, N_CELL)
      call dump_i4('No_cp', No_cp, 3*N_CELL)
      call dump_r8('dxVV', dxVV, 3*N_CELL0)
      call dump_r8('dyVV', dyVV, 3*N_CELL0)
      call dump_r8('dlVV', dlVV, 3*N_CELL0)
      call dump_r8('xnn', xnn, 3*N_CELL0)
      call dump_r8('ynn', ynn, 3*N_CELL0)
      call dump_r8('dxCC', dxCC, 3*N_CELL0)
      call dump_r8('dyCC', dyCC, 3*N_CELL0)
      call dump_r8('dlCC', dlCC, 3*N_CELL0)
      call dump_r8('xme', xme, 3*N_CELL0)
      call dump_r8('yme', yme, 3*N_CELL0)
      call dump_r8('xmo', xmo, 3*N_CELL0)
      call dump_r8('ymo', ymo, 3*N_CELL0)
      call dump_r8('doe', doe, 3*N_CELL0)
      call dump_r8('dle1', dle1, 3*N_CELL0)
      call dump_r8('dle2', dle2, 3*N_CELL0)
      call dump_r8('dle3', dle3, 3*N_CELL0)
      call dump_r8('sum_xc2', sum_xc2, N_CELL0)
      call dump_r8('sum_yc2', sum_yc2, N_CELL0)
      call dump_r8('sum_xcyc', sum_xcyc, N_CELL0)
      call dump_i4('surrounding', surrounding, 10*N_VERT)
      call dump_r8('weight', weight, 10*N_VERT)
      call dump_i4('Dimsurrounding', Dimsurrounding, N_VERT)
      call dump_r8('dlVsum', dlVsum, N_VERT)
      call dump_i4('vertex_pair', vertex_pair, N_VERT)
      call dump_i4('vertex_edge', vertex_edge, N_VERT)
      call dump_i4('tagXV', tagXV, N_VERT)
      call dump_i4('tagYV', tagYV, N_VERT)
      call dump_i4('bcuc', bcuc, N_CELL)
      call dump_i4('bcvc', bcvc, N_CELL)
      call dump_i4('bcwc', bcwc, N_CELL)
      call dump_i4('bcpc', bcpc, N_CELL)
      call dump_i4('bcuv', bcuv, N_VERT)
      call dump_i4('bcvv', bcvv, N_VERT)
      call dump_i4('bcwv', bcwv, N_VERT)
      call dump_i4('bcpv', bcpv, N_VERT)
      call dump_r8('AmC0', AmC0, N_CELL0*NZ)
      call dump_r8('AmC1', AmC1, N_CELL0*NZ)
      call dump_r8('AmC2', AmC2, N_CELL0*NZ)
      call dump_r8('AmC3', AmC3, N_CELL0*NZ)
      call dump_r8('AmCT', AmCT, N_CELL0*NZ)
      call dump_r8('AmCB', AmCB, N_CELL0*NZ)
      call dump_r8('sig', sig, NZ)
      call dump_r8('dsig', dsig, NZ)
      call dump_r8('sigv', sigv, NZ-1)
      call dump_r8('dsigv', dsigv, NZ-1)
!     the Tecplot cell-centre elements (entries past Numictec are unset)
      do n1 = Numictec+1, 5*N_VERT
         ic1tec(n1) = 0
         ic2tec(n1) = 0
         ic3tec(n1) = 0
      enddo
      call dump_i4('ic1tec', ic1tec, 5*N_VERT)
      call dump_i4('ic2tec', ic2tec, 5*N_VERT)
      call dump_i4('ic3tec', ic3tec, 5*N_VERT)
      iv(1) = Numictec
      call dump_i4('Numictec', iv, 1)
      end subroutine

      subroutine dump_state(sfx)
      character(*) :: sfx
      call dump_r8('ufn'//sfx, ufn, N_CELL*NZ)
      call dump_r8('vfn'//sfx, vfn, N_CELL*NZ)
      call dump_r8('wfn'//sfx, wfn, N_CELL*NZ)
      call dump_r8('pfn'//sfx, pfn, N_CELL*NZ)
      call dump_r8('ufv'//sfx, ufv, N_VERT*(NZ-1))
      call dump_r8('vfv'//sfx, vfv, N_VERT*(NZ-1))
      call dump_r8('wfv'//sfx, wfv, N_VERT*(NZ-1))
      call dump_r8('pfv'//sfx, pfv, N_VERT*(NZ-1))
      call dump_r8('U1FACE'//sfx, U1FACE, N_CELL0*NZ)
      call dump_r8('U2FACE'//sfx, U2FACE, N_CELL0*NZ)
      call dump_r8('U3FACE'//sfx, U3FACE, N_CELL0*NZ)
      call dump_r8('UTFACE'//sfx, UTFACE, N_CELL0*NZ)
      call dump_r8('UBFACE'//sfx, UBFACE, N_CELL0*NZ)
      call dump_r8('rhsuf'//sfx, rhsuf, N_CELL*NZ)
      call dump_r8('rhsvf'//sfx, rhsvf, N_CELL*NZ)
      call dump_r8('rhswf'//sfx, rhswf, N_CELL*NZ)
      if (allocated(eddy)) call dump_r8('eddy'//sfx, eddy, N_CELL*NZ)
      end subroutine

!     single routines on the initial field; inputs are copied first so that
!     each call sees the same state
!     mode 3: the set-up alone, with the mesh as read
      subroutine setup_mode
      call dump_geometry
      call dump_i4('mesh_No_vp', No_vp, 3*N_CELL0)
      call dump_i4('mesh_nbe', nbe, N_CELL0)
      call dump_i4('mesh_nbev', nbev, N_VERT)
      call dump_r8('mesh_xv', xv, N_VERT)
      call dump_r8('mesh_yv', yv, N_VERT)
      call dump_r8('mesh_zbv', zbv, N_VERT)
      end subroutine

      subroutine unit_mode
      integer :: m1, m2
      call dump_geometry
      call dump_state('_0')
      allocate(w1(N_CELL,NZ),w2(N_CELL,NZ),w3(N_CELL,NZ),w4(N_CELL,NZ),w5(N_CELL,NZ),w6(N_CELL,NZ))
      allocate(g1(N_CELL,NZ),g2(N_CELL,NZ),g3(N_CELL,NZ),t1(N_CELL,NZ),t2(N_CELL,NZ),t3(N_CELL,NZ))
      allocate(v1(N_VERT,NZ-1),v2(N_VERT,NZ-1),v3(N_VERT,NZ-1),v4(N_VERT,NZ-1))
      allocate(f1(N_CELL0,NZ),f2(N_CELL0,NZ),f3(N_CELL0,NZ),f4(N_CELL0,NZ),f5(N_CELL0,NZ))

!     cell -> vertex, tagdir 2 and 3
      w1 = ufn
      w2 = vfn
      w3 = wfn
      v1 = 0.0d0
      v2 = 0.0d0
      v3 = 0.0d0
      call BCglobalVC(w1,w2,w3,v1,v2,v3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev,2)
      call dump_r8('vc2_u', v1, N_VERT*(NZ-1))
      call dump_r8('vc2_v', v2, N_VERT*(NZ-1))
      call dump_r8('vc2_w', v3, N_VERT*(NZ-1))
      w4 = pfn
      v4 = 0.0d0
      call BCglobalP(w4,v4, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev,2)
      call dump_r8('pv2', v4, N_VERT*(NZ-1))
      v4 = 0.0d0
      call BCglobalP(w4,v4, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev,3)
      call dump_r8('pv3', v4, N_VERT*(NZ-1))

!     least-squares gradient of p and of u
      call grandientLSM(g1,g2,g3,w4,xc,yc,sig,dsig,No_cp,nbe)
      call dump_r8('lsm_p_x', g1, N_CELL*NZ)
      call dump_r8('lsm_p_y', g2, N_CELL*NZ)
      call dump_r8('lsm_p_s', g3, N_CELL*NZ)
      call grandientLSM(t1,t2,t3,w1,xc,yc,sig,dsig,No_cp,nbe)
      call dump_r8('lsm_u_x', t1, N_CELL*NZ)
      call dump_r8('lsm_u_y', t2, N_CELL*NZ)
      call dump_r8('lsm_u_s', t3, N_CELL*NZ)

!     face values of u (Kim-Choi)
      f1 = 0.0d0
      f2 = 0.0d0
      f3 = 0.0d0
      f4 = 0.0d0
      f5 = 0.0d0
      call calcul_face(f1,f2,f3,f4,f5,w1,v1, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev,2)
      call dump_r8('face_u_1', f1, N_CELL0*NZ)
      call dump_r8('face_u_2', f2, N_CELL0*NZ)
      call dump_r8('face_u_3', f3, N_CELL0*NZ)
      call dump_r8('face_u_T', f4, N_CELL0*NZ)
      call dump_r8('face_u_B', f5, N_CELL0*NZ)

!     Smagorinsky eddy viscosity
      w1 = ufn
      w2 = vfn
      w3 = wfn
      call sgles(w1,w2,w3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev, &
                 Hpr,h,eta,etan, Hprv,hv,etav)
      call dump_r8('eddy_0', eddy, N_CELL*NZ)

!     momentum fluxes (AdvDiffNS) with the forcing as starting rhs
      do m2=1,NZ
         do m1=1,N_CELL
            g1(m1,m2) = bdx
            g2(m1,m2) = bdy
            g3(m1,m2) = bdz
            t1(m1,m2) = eddy(m1,m2)+1./Re
         enddo
      enddo
      w1 = ufn
      w2 = vfn
      w3 = wfn
      call AdvDiffNS(g1,g2,g3, w1,w2,w3, t1,t1,t1, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev)
      call dump_r8('rhsu', g1, N_CELL*NZ)
      call dump_r8('rhsv', g2, N_CELL*NZ)
      call dump_r8('rhsw', g3, N_CELL*NZ)

!     divergence source and its Kim-Choi correction
      w1 = ufn
      w2 = vfn
      w3 = wfn
      call source3D(w5,w1,w2,w3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev)
      call dump_r8('src', w5, N_CELL*NZ)
      call dump_r8('src_U1', U1FACE, N_CELL0*NZ)
      call dump_r8('src_U2', U2FACE, N_CELL0*NZ)
      call dump_r8('src_U3', U3FACE, N_CELL0*NZ)
      call dump_r8('src_UT', UTFACE, N_CELL0*NZ)
      call dump_r8('src_UB', UBFACE, N_CELL0*NZ)
      w4 = pfn
      v4 = 0.0d0
      call gradientP(w4,v4,t1,t2,t3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev)
      do m2=2,NZ-1
         do m1=1,N_CELL0
            w5(m1,m2) = w5(m1,m2)/dt
         enddo
      enddo
      call update_rhsp(w5,w4,v4,t1,t2,t3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev)
      call dump_r8('rhsp', w5, N_CELL*NZ)

!     pressure solve from p = pfn
      w6 = pfn
      v4 = 0.0d0
      call bicgstab(w6,v4,w5, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev, &
                    Hpr,h,etan, Hprv,hv,etav)
      call dump_r8('psol', w6, N_CELL*NZ)
      call dump_r8('psol_v', v4, N_VERT*(NZ-1))

!     face-velocity correction with the solved pressure
      call gradientP(w6,v4,t1,t2,t3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev)
      call dump_r8('gp_x', t1, N_CELL*NZ)
      call dump_r8('gp_y', t2, N_CELL*NZ)
      call dump_r8('gp_s', t3, N_CELL*NZ)
      call newu3D(w6,v4,flag_ab,t1,t2,t3, xc,yc,sig,dsig,No_cp,nbe, xv,yv,sigv,dsigv,No_vp,nbev)
      call dump_r8('newu_U1', U1FACE, N_CELL0*NZ)
      call dump_r8('newu_U2', U2FACE, N_CELL0*NZ)
      call dump_r8('newu_U3', U3FACE, N_CELL0*NZ)
      call dump_r8('newu_UT', UTFACE, N_CELL0*NZ)
      call dump_r8('newu_UB', UBFACE, N_CELL0*NZ)
      call dump_r8('newu_pv', v4, N_VERT*(NZ-1))
      end subroutine

!     time loop of nsmp3D.F90:396-564 (FinRK stages per step, RK2 average)
      subroutine step_mode
      character(len=8) :: sfx
      integer :: m1, m2
      call dump_geometry
      call dump_state('_0')
      call LoopTimeInitial(etanp,                              &
                           alphafnp,ufnp,vfnp,wfnp,pfnp,       &
                           alphasnp,usnp,vsnp,wsnp,psnp,       &
                           xct,yct,zct,                        &
                           xvt,yvt,zvt,                        &
                           etan,                               &
                           alphafn,ufn,vfn,wfn,pfn,            &
                           alphasn,usn,vsn,wsn,psn,            &
                           xc,yc,sig,Hpr,h,                    &
                           xv,yv,sigv,Hprv,hv,                 &
                           No_cp,nbe)
      flag_ab = 0
      call alloc_stats_variables
      call sm_init()
      do istep=1,nsteps
         time  = time + dt
         nstep = nstep + 1
         print*, 'NSX_STEP', istep
         do m1=1,N_CELL
            eta(m1) = etan(m1)
            do m2=1,NZ
               alphaf(m1,m2) = alphafn(m1,m2)
               alphas(m1,m2) = alphasn(m1,m2)
               uf(m1,m2) = ufn(m1,m2)
               vf(m1,m2) = vfn(m1,m2)
               wf(m1,m2) = wfn(m1,m2)
               us(m1,m2) = usn(m1,m2)
               vs(m1,m2) = vsn(m1,m2)
               ws(m1,m2) = wsn(m1,m2)
               pf(m1,m2) = pfn(m1,m2)
               ps(m1,m2) = psn(m1,m2)
            enddo
         enddo
         do irk=1,FinRK
            print*, 'NSX_STAGE', irk
            call hydro(ufnp,ufn,uf,ufv,                          &
                       vfnp,vfn,vf,vfv,                          &
                       wfnp,wfn,wf,wfv,                          &
                       pfnp,pfn,pf,pfv,                          &
                       rhof,viscof,                              &
                       rhofv,viscofv,                            &
                       xc,yc,sig,dsig,No_cp,nbe,                 &
                       xv,yv,sigv,dsigv,No_vp,nbev,              &
                       Hpr,h,etanp,etan,                         &
                       Hprv,hv,etav,                             &
                       flag_ab)
            if (istep .eq. 1 .and. irk .eq. 1) then
               call dump_r8('s1_ufnp', ufnp, N_CELL*NZ)
               call dump_r8('s1_vfnp', vfnp, N_CELL*NZ)
               call dump_r8('s1_wfnp', wfnp, N_CELL*NZ)
               call dump_r8('s1_pfnp', pfnp, N_CELL*NZ)
               call dump_r8('s1_ufv', ufv, N_VERT*(NZ-1))
               call dump_r8('s1_pfv', pfv, N_VERT*(NZ-1))
               call dump_r8('s1_U1', U1FACE, N_CELL0*NZ)
               call dump_r8('s1_UT', UTFACE, N_CELL0*NZ)
               call dump_r8('s1_rhsuf', rhsuf, N_CELL*NZ)
               call dump_r8('s1_eddy', eddy, N_CELL*NZ)
            endif
            if (FinRK .eq. 2) then
               do m1=1,N_CELL
                  eta(m1) = etanp(m1)
                  do m2=1,NZ
                     alphaf(m1,m2) = alphafnp(m1,m2)
                     alphas(m1,m2) = alphasnp(m1,m2)
                     uf(m1,m2) = ufnp(m1,m2)
                     vf(m1,m2) = vfnp(m1,m2)
                     wf(m1,m2) = wfnp(m1,m2)
                     us(m1,m2) = usnp(m1,m2)
                     vs(m1,m2) = vsnp(m1,m2)
                     ws(m1,m2) = wsnp(m1,m2)
                     pf(m1,m2) = pfnp(m1,m2)
                     ps(m1,m2) = psnp(m1,m2)
                  enddo
               enddo
            endif
         enddo
         if (FinRK .ne. 1) then
            do m1=1,N_CELL
               do m2=1,NZ
                  alphafnp(m1,m2) = 0.5d0*(alphafn(m1,m2)+alphafnp(m1,m2))
                  alphasnp(m1,m2) = 0.5d0*(alphasn(m1,m2)+alphasnp(m1,m2))
                  ufnp(m1,m2) = 0.5d0*(ufn(m1,m2)+ufnp(m1,m2))
                  vfnp(m1,m2) = 0.5d0*(vfn(m1,m2)+vfnp(m1,m2))
                  wfnp(m1,m2) = 0.5d0*(wfn(m1,m2)+wfnp(m1,m2))
                  usnp(m1,m2) = 0.5d0*(usn(m1,m2)+usnp(m1,m2))
                  vsnp(m1,m2) = 0.5d0*(vsn(m1,m2)+vsnp(m1,m2))
                  wsnp(m1,m2) = 0.5d0*(wsn(m1,m2)+wsnp(m1,m2))
               enddo
            enddo
         endif
         call LoopTimeUpdate(etan,                                &
                             alphafn,ufn,vfn,wfn,pfn,             &
                             alphasn,usn,vsn,wsn,psn,             &
                             xct,yct,zct,                         &
                             xvt,yvt,zvt,                         &
                             etanp,                               &
                             alphafnp,ufnp,vfnp,wfnp,pfnp,        &
                             alphasnp,usnp,vsnp,wsnp,psnp,        &
                             xc,yc,sig,Hpr,h,                     &
                             xv,yv,sigv,Hprv,hv,                  &
                             No_cp,nbe)
         if (mod(istep, hdr(10)) .eq. 0 .or. istep .eq. 1 .or. istep .eq. nsteps) then
            write(sfx,'(a,i0)') '_', istep
            call dump_state(trim(sfx))
         endif
         call glob_sample(ufnp,vfnp,wfnp,pfnp)
      enddo
!     the CFL / mass-flux diagnostics of nsmp3D.F90:595-606 on the final state
      print*, 'NSX_CFL'
      call glob_cfl(ufn,vfn,wfn,pfn,xc,yc,sig,dsig,No_cp,nbe)
      call glob_put_plane_mm()
      call dump_r8('sm_uf', sm_uf, N_CELL*NZ)
      call dump_r8('sm_vf', sm_vf, N_CELL*NZ)
      call dump_r8('sm_wf', sm_wf, N_CELL*NZ)
      call dump_r8('sm_u2f', sm_u2f, N_CELL*NZ)
      call dump_r8('sm_v2f', sm_v2f, N_CELL*NZ)
      call dump_r8('sm_w2f', sm_w2f, N_CELL*NZ)
      call dump_r8('sm_uvf', sm_uvf, N_CELL*NZ)
      call dump_r8('sm_vwf', sm_vwf, N_CELL*NZ)
      call dump_r8('sm_wuf', sm_wuf, N_CELL*NZ)
      call dump_r8('sm_pf', sm_pf, N_CELL*NZ)
      call dump_r8('sm_p2f', sm_p2f, N_CELL*NZ)
      call dump_r8('sm_ef', sm_ef, N_CELL*NZ)
      SaveCounter = 3
      call SavetecVertex(alphafv,ufv,vfv,wfv,pfv,           &
                         alphasv,usv,vsv,wsv,psv,rhosv,     &
                         xvt,yvt,zvt,No_vp)
      call SavetecCenter(alphafnp,ufnp,vfnp,wfnp,pfnp,      &
                         alphasnp,usnp,vsnp,wsnp,psnp,rhos, &
                         xvt,yvt,zvt,No_vp)
      call SaveParaviewVertex(alphafv,ufv,vfv,wfv,pfv,      &
                              alphasv,usv,vsv,wsv,psv,rhosv,&
                              xvt,yvt,zvt,No_vp)
      call SavetecVC(alphafnp,ufnp,vfnp,wfnp,pfnp,          &
                     alphasnp,usnp,vsnp,wsnp,psnp,rhos,     &
                     xct,yct,zct,No_cp,                     &
                     alphafv,ufv,vfv,wfv,pfv,               &
                     alphasv,usv,vsv,wsv,psv,rhosv,         &
                     xvt,yvt,zvt,No_vp)
      call dump_r8('usv', usv, N_VERT*(NZ-1))
      call dump_r8('vsv', vsv, N_VERT*(NZ-1))
      call dump_r8('wsv', wsv, N_VERT*(NZ-1))
      call dump_r8('xct', xct, N_CELL*NZ)
      call dump_r8('yct', yct, N_CELL*NZ)
      call dump_r8('zct', zct, N_CELL*NZ)
      call dump_r8('xvt', xvt, N_VERT*(NZ-1))
      call dump_r8('yvt', yvt, N_VERT*(NZ-1))
      call dump_r8('zvt', zvt, N_VERT*(NZ-1))
      tv(1) = time
      call dump_r8('time', tv, 1)
      end subroutine

      end program
