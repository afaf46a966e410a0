! Driver program for the MI355X lattice-update library, keeping the reference driver's inputs
! and files: infile.in in the line order of src/m_readinfile.F90:84-160, restart files
! restart/restart_<tile4>_<it6>.uf (src/m_saverestart.F90:125-136) and
! restart/turbulence_<tile4>_<it6>.uf, the end-of-run testing file (src/m_testing.F90), and the
! flow of src/main.F90: read inputs, restart or equilibrium start, time loop, final restart.
!
! The grid comes from the restart header when nt0 > 0 (the reference fixes it at compile time in
! mod_dimensions); for nt0 = 0 give it on the command line:  lbm_driver nx ny nz [nl].
! Inflow-turbulence fields come from the turbulence restart file when it exists, otherwise (and
! every nrturb steps, as src/main.F90:381-383) from the device generator lbm_generate_turbulence.
! Time averages (laveraging, avestart, avesave) are accumulated on the device and written to
! averages/averages_0000_<avesave>.uf.  Actuator-line turbines (the infile's turbine section, the
! NREL 5-MW blade, ./Airfoils/*.dat, restart/theta_0000_<it6>.uf) run on the device.  Not carried
! over (outside the lattice update): the transport of the potential temperature that iablvisc = 2
! needs (heatflux + advection; the library's buoyancy forcing, lbm_set_buoyancy, takes the field
! from its caller), tracers, Tecplot and NetCDF output, the random
! perturbation of inipert (nt0 = 0 starts from the inflow profile).  Geometry: 'none' or the
! reference's 'cylinder' (src/m_cylinder.F90).
program lbm_driver
   use, intrinsic :: iso_c_binding
   use lbm_hip
   implicit none
#ifdef DOUBLE_PRECISION
   integer, parameter :: rk = c_double
#else
   integer, parameter :: rk = c_float
#endif
   integer(c_int), parameter :: realkind = rk

   ! infile.in
   logical :: ltiming, ltesting, ldump, lmeasurements, inflowturbulence, laveraging, laveturb
   integer :: ntx, nty, ntz, ibgk, ihrr, ivreman, iablvisc, istable, nt0, nt1, iout, irestart
   integer :: iprt1, iprt2, dprt, itecout, ibnd, jbnd, kbnd, nrturb, avestart, avesave, nturbines
   real(rk) :: smagorinsky, ablheight, uini, udir, turbulence_ampl, visc, p_rho, p_len, p_vel
   real(rk) :: p_time, p_visc, kinevisc
   character(len=20) :: experiment
   character(len=1) :: ver

   integer :: nx, ny, nz, nl, i, j, k, it, iunit, nchunk, nargs, ios
   character(len=32) :: arg
   character(len=100) :: fname
   real(rk), allocatable, target :: f(:,:,:,:), rho(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:)
   real(rk), allocatable, target :: uu(:,:,:), vv(:,:,:), ww(:,:,:), rr(:,:,:)
   real(rk), allocatable, target :: uvel_shear(:), uvel_time(:), udir_time(:)
   integer(c_int8_t), allocatable, target :: mask(:,:,:)
   logical :: ex, lsolids
   type(lbm_desc) :: desc
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   integer(8) :: t0, t1, rate
   real(8) :: wall, mass, mom(3)
   ! turbines: infile values (src/m_readinfile.F90:143-161) and the device description
   real(rk) :: pitchangle, turbrpm, tipspeedratio
   integer :: itiploss, nang_max
   integer, allocatable :: ipos(:), jpos(:), kpos(:)
   real(rk), allocatable :: yawd(:), tiltd(:)
   real(c_double), allocatable, target :: t_hub(:,:), t_yaw(:), t_tilt(:), t_omegand(:), t_pitch(:), t_theta(:)
   real(c_double), allocatable, target :: t_relm(:), t_dc(:), t_chord(:), t_twist(:), t_foil(:,:,:)
   integer(c_int32_t), allocatable, target :: t_nfoil(:), t_nang(:)
   type(lbm_turbines) :: tdesc
   ! NREL 5-MW blade (src/mod_nrel5mw.F90:24-48) and its airfoil tables (src/m_nrelreadfoil.F90:13-20)
   real(rk), parameter :: nrel_relm(17) = [2.86670_rk, 5.60000_rk, 8.33330_rk, 11.75000_rk, 15.85000_rk, &
      19.95000_rk, 24.05000_rk, 28.15000_rk, 32.25000_rk, 36.35000_rk, 40.45000_rk, 44.55000_rk, 48.65000_rk, &
      52.75000_rk, 56.16670_rk, 58.90000_rk, 61.63330_rk]
   real(rk), parameter :: nrel_dc(17) = [2.73330_rk, 2.73330_rk, 2.73330_rk, 4.10000_rk, 4.10000_rk, 4.10000_rk, &
      4.10000_rk, 4.10000_rk, 4.10000_rk, 4.10000_rk, 4.10000_rk, 4.10000_rk, 4.10000_rk, 4.10000_rk, 2.73330_rk, &
      2.73330_rk, 2.73330_rk]
   real(rk), parameter :: nrel_chord(17) = [3.54200_rk, 3.85400_rk, 4.16700_rk, 4.55700_rk, 4.65200_rk, 4.45800_rk, &
      4.24900_rk, 4.00700_rk, 3.74800_rk, 3.50200_rk, 3.25600_rk, 3.01000_rk, 2.76400_rk, 2.51800_rk, 2.31300_rk, &
      2.08600_rk, 1.41900_rk]
   real(rk), parameter :: nrel_twist(17) = [13.3080_rk, 13.3080_rk, 13.3080_rk, 13.3080_rk, 11.4800_rk, 10.1620_rk, &
      9.0110_rk, 7.7950_rk, 6.5440_rk, 5.3610_rk, 4.1880_rk, 3.1250_rk, 2.3190_rk, 1.5260_rk, 0.8630_rk, 0.3700_rk, &
      0.1060_rk]
   integer, parameter :: nrel_nfoil(17) = [1, 1, 2, 3, 4, 4, 5, 6, 6, 7, 7, 8, 8, 8, 8, 8, 8]
   character(len=14), parameter :: foil_files(8) = [character(len=14) :: 'Cylinder1.dat', 'Cylinder2.dat', &
      'DU40_A17.dat', 'DU35_A17.dat', 'DU30_A17.dat', 'DU25_A17.dat', 'DU21_A17.dat', 'NACA64_A17.dat']

   ! ---------------------------------------------------------------- infile.in
   inquire(file='infile.in', exist=ex)
   if (.not. ex) stop 'lbm_driver: infile.in not found'
   open(newunit=iunit, file='infile.in', status='old')
   read(iunit,'(a)') ver
   read(iunit,*) ltiming
   read(iunit,*) ltesting
   read(iunit,*) ldump
   read(iunit,*) ntx, nty, ntz
   read(iunit,*) lmeasurements
   read(iunit,'(a)') ver
   read(iunit,*) experiment
   read(iunit,*) ibgk
   read(iunit,*) ihrr
   read(iunit,*) ivreman, smagorinsky
   read(iunit,'(a)') ver
   read(iunit,*) iablvisc
   read(iunit,*) ablheight
   read(iunit,*) istable
   read(iunit,'(a)') ver
   read(iunit,*) nt0
   read(iunit,*) nt1
   if (nt1 <= nt0) stop 'lbm_driver: nt1 <= nt0'
   read(iunit,*) iout
   read(iunit,*) irestart
   read(iunit,*) iprt1, iprt2, dprt
   read(iunit,*) itecout
   read(iunit,'(a)') ver
   read(iunit,*) ibnd
   read(iunit,*) jbnd
   read(iunit,*) kbnd
   read(iunit,'(a)') ver
   read(iunit,*) uini, udir
   read(iunit,*) inflowturbulence, turbulence_ampl, nrturb
   read(iunit,'(a)') ver
   read(iunit,*) visc
   read(iunit,*) p_rho
   read(iunit,*) p_len
   read(iunit,*) p_vel
   read(iunit,'(a)') ver
   read(iunit,'(1x,l1,1x,l1)') laveraging, laveturb
   read(iunit,*) avestart
   read(iunit,*) avesave
   read(iunit,'(a)') ver
   read(iunit,*) nturbines
   if (nturbines > 0) then
      allocate(ipos(nturbines), jpos(nturbines), kpos(nturbines), yawd(nturbines), tiltd(nturbines))
      read(iunit,*) pitchangle
      read(iunit,*) turbrpm
      read(iunit,*) tipspeedratio
      read(iunit,*) itiploss
      do i = 1, nturbines
         read(iunit,'(a)') ver
         read(iunit,*) ipos(i), jpos(i), kpos(i), yawd(i), tiltd(i)
      enddo
   endif
   close(iunit)
   if (iablvisc == 2) stop 'lbm_driver: iablvisc = 2 needs the potential-temperature transport, not in this driver'
   ! conversions as src/m_readinfile.F90:133,171-186
   uini = uini / p_vel
   p_time = p_len / p_vel
   p_visc = p_len**2 / p_time
   kinevisc = visc / p_visc

   ! ---------------------------------------------------------------- grid and state
   nl = 27
   if (nt0 > 0) then
      write(fname,'(a,i6.6,a)') 'restart/restart_0000_', nt0, '.uf'
      inquire(file=trim(fname), exist=ex)
      if (.not. ex) then
         print '(2a)', 'lbm_driver: restart file does not exist: ', trim(fname)
         stop 1
      endif
      open(newunit=iunit, file=trim(fname), form='unformatted', status='old')
      read(iunit) nx, ny, nz, nl
      close(iunit)
   else
      nargs = command_argument_count()
      if (nargs < 3) stop 'lbm_driver: nt0 = 0 needs the grid on the command line: nx ny nz [nl]'
      call get_command_argument(1, arg); read(arg,*) nx
      call get_command_argument(2, arg); read(arg,*) ny
      call get_command_argument(3, arg); read(arg,*) nz
      if (nargs >= 4) then
         call get_command_argument(4, arg); read(arg,*) nl
      endif
   endif
   print '(a,3i6,a,i3)', 'lbm_driver: grid ', nx, ny, nz, '  nl=', nl

   desc%nx = nx; desc%ny = ny; desc%nz = nz; desc%q = nl
   desc%storage = merge(LBM_STORE_F64, LBM_STORE_F32, rk == c_double)
   desc%ibgk = ibgk; desc%ihrr = ihrr; desc%ivreman = ivreman
   desc%ibnd = ibnd; desc%jbnd = jbnd; desc%kbnd = kbnd
   desc%iturb_pos = 10; desc%store_tau = 0; desc%block_x = 0; desc%block_y = 0
   desc%kinevisc = real(kinevisc, c_double)
   desc%smagorinsky = real(smagorinsky, c_double)
   desc%rho0 = 1.0d0
   rc = lbm_create(desc, ctx)
   call lbm_check(c_null_ptr, rc, 'lbm_create')

   allocate(f(nl,0:nx+1,0:ny+1,0:nz+1))
   allocate(rho(0:nx+1,0:ny+1,0:nz+1), u(0:nx+1,0:ny+1,0:nz+1), v(0:nx+1,0:ny+1,0:nz+1), w(0:nx+1,0:ny+1,0:nz+1))

   ! solids (src/m_solid_objects_init.F90:43-58)
   allocate(mask(0:nx+1,0:ny+1,0:nz+1))
   mask = 0
   lsolids = .false.
   select case (trim(experiment))
   case ('cylinder')
      ! radius 7 at i = 50, j = ny/2, every k (src/m_cylinder.F90:8-17)
      do j = 1, ny
         do i = 1, nx
            if ((i - 50)**2 + (j - ny/2)**2 < 49) mask(i, j, :) = 1
         enddo
      enddo
      lsolids = .true.
   case ('none', '')
   case default
      print '(2a)', 'lbm_driver: geometry not part of this build: ', trim(experiment)
      stop 1
   end select
   if (lsolids) then
      rc = lbm_set_blanking(ctx, c_loc(mask))
      call lbm_check(ctx, rc, 'lbm_set_blanking')
   endif

   ! inflow profile and time series (src/m_inflow.F90)
   allocate(uvel_shear(nz), uvel_time(nt0:nt1), udir_time(nt0:nt1))
   call read_inflow()
   rc = lbm_set_inflow(ctx, c_loc(uvel_shear), real(uini, c_double), real(udir, c_double), realkind)
   call lbm_check(ctx, rc, 'lbm_set_inflow')
   rc = lbm_set_inflow_series(ctx, nt0, nt1, c_loc(uvel_time), c_loc(udir_time), realkind)
   call lbm_check(ctx, rc, 'lbm_set_inflow_series')

   if (nt0 == 0) then
      ! equilibrium start from the inflow profile (src/main.F90:214,243-247; no random part)
      rho = 1.0_rk
      do k = 0, nz + 1
         u(:,:,k) = uvel_time(nt0) * uvel_shear(min(max(k,1),nz)) * cos(udir_time(nt0) * 3.1415927410125732_rk / 180.0_rk)
         v(:,:,k) = uvel_time(nt0) * uvel_shear(min(max(k,1),nz)) * sin(udir_time(nt0) * 3.1415927410125732_rk / 180.0_rk)
      enddo
      w = 0.0_rk
      rc = lbm_init_equilibrium(ctx, c_loc(rho), c_loc(u), c_loc(v), c_loc(w), realkind, nt0)
      call lbm_check(ctx, rc, 'lbm_init_equilibrium')
   else
      ! restart (src/m_readrestart.F90:163-185)
      write(fname,'(a,i6.6,a)') 'restart/restart_0000_', nt0, '.uf'
      open(newunit=iunit, file=trim(fname), form='unformatted', status='old')
      read(iunit) i, j, k, nl, f
      close(iunit)
      rc = lbm_set_state(ctx, c_loc(f), realkind)
      call lbm_check(ctx, rc, 'lbm_set_state')
   endif

   if (nturbines > 0) call setup_turbines()

   if (inflowturbulence) then
      allocate(uu(ny,nz,0:nrturb), vv(ny,nz,0:nrturb), ww(ny,nz,0:nrturb), rr(ny,nz,0:nrturb))
      write(fname,'(a,i6.6,a)') 'restart/turbulence_0000_', nt0, '.uf'
      inquire(file=trim(fname), exist=ex)
      if (nt0 == 0 .or. .not. ex) then
         ! generated on the device (src/main.F90:252, inflow_turbulence_update)
         call generate_turbulence(nt0)
      else
         ! turbulence fields from restart/turbulence_0000_<nt0>.uf (src/m_readrestart.F90:66-100)
         open(newunit=iunit, file=trim(fname), form='unformatted', status='old')
         read(iunit) j, k, i
         rewind(iunit)
         if (j /= ny .or. k /= nz .or. i /= nrturb) stop 'lbm_driver: incompatible turbulence restart file'
         read(iunit) j, k, i, uu, vv, ww, rr
         close(iunit)
         rc = lbm_set_turbulence(ctx, c_loc(uu), c_loc(vv), c_loc(ww), int(nrturb, c_int), &
                                 real(turbulence_ampl, c_double), realkind)
         call lbm_check(ctx, rc, 'lbm_set_turbulence')
      endif
   endif

   ! ---------------------------------------------------------------- time loop
   print '(a)', ' Start timestepping loop'
   call system_clock(t0, rate)
   it = nt0
   do while (it < nt1)
      nchunk = min(10 - mod(it, 10), nt1 - it)
      if (ldump .and. irestart > 0) nchunk = min(nchunk, irestart - mod(it, irestart))
      if (inflowturbulence .and. ibnd == 1) nchunk = min(nchunk, nrturb - mod(it, nrturb))
      if (laveraging .and. it < avestart) nchunk = min(nchunk, avestart - it)
      if (laveraging .and. it < avesave) nchunk = min(nchunk, avesave - it)
      rc = lbm_step(ctx, int(it, c_int), int(nchunk, c_int))
      call lbm_check(ctx, rc, 'lbm_step')
      it = it + nchunk
      if (mod(it, 10) == 0 .or. it == nt1) print '(a,i6,a,f10.2,a)', 'Iteration: ', it, ' Time: ', real(it)*p_time, ' s.'
      ! time averaging over steps avestart+1 .. avesave (src/main.F90:374-377)
      if (laveraging .and. it == avestart) then
         rc = lbm_set_averaging(ctx, 1_c_int)
         call lbm_check(ctx, rc, 'lbm_set_averaging')
      endif
      if (laveraging .and. it == avesave) call save_averages(it)
      ! new turbulence fields every nrturb steps (src/main.F90:381-383)
      if (inflowturbulence .and. ibnd == 1 .and. mod(it, nrturb) == 0 .and. it > 1 .and. it < nt1) &
         call generate_turbulence(it)
      if (ldump .and. irestart > 0 .and. mod(it, irestart) == 0 .and. it /= nt1) call save_restart(it)
   enddo
   rc = lbm_synchronize(ctx)
   call lbm_check(ctx, rc, 'lbm_synchronize')
   call system_clock(t1)
   wall = real(t1 - t0, 8) / real(rate, 8)
   rc = lbm_get_totals(ctx, mass, mom)
   call lbm_check(ctx, rc, 'lbm_get_totals')
   print '(a,es16.8,a,3es14.6)', 'mass =', mass, '  momentum =', mom
   if (ltiming) then
      print '(a,1x,f13.5)', 'Total TS routines  time =', wall
      print '(a,1x,f13.2)', 'MLUPS                   =', real(nx,8)*ny*nz*(nt1 - nt0) / wall / 1.0d6
   endif

   if (ldump) call save_restart(nt1)
   call testing(nt1)
   rc = lbm_destroy(ctx)

contains

   subroutine setup_turbines()
      ! turbine_initialize (src/m_turbine_initialize.F90:29-70) in the REAL kind, the airfoil tables
      ! (src/m_nrelreadfoil.F90:33-47) and the azimuths of the restart (src/m_readrestart.F90:103-116)
      real(rk), parameter :: pi = 3.14159265358979323846_rk, pi2 = 2.0_rk * pi
      real(rk) :: omega, row(4)
      real(rk), allocatable :: th(:)
      integer :: iu, kf, ia, na(8)
      character(len=100) :: fn
      omega = pi2 * turbrpm / 60.0_rk
      allocate(t_hub(3, nturbines), t_yaw(nturbines), t_tilt(nturbines), t_omegand(nturbines), t_pitch(nturbines), &
               t_theta(nturbines), th(nturbines))
      do i = 1, nturbines
         t_hub(:, i) = [real(real(ipos(i), rk), c_double), real(real(jpos(i), rk), c_double), &
                        real(real(kpos(i), rk), c_double)]
         t_yaw(i) = real((yawd(i) / 360.0_rk) * pi2, c_double)
         t_tilt(i) = real((tiltd(i) / 360.0_rk) * pi2, c_double)
         t_omegand(i) = real(omega * p_time, c_double)
         t_pitch(i) = real(pitchangle, c_double)
      enddo
      allocate(t_relm(17), t_dc(17), t_chord(17), t_twist(17), t_nfoil(17))
      t_relm = real(nrel_relm / p_len, c_double)
      t_dc = real(nrel_dc / p_len, c_double)
      t_chord = real(nrel_chord / p_len, c_double)
      t_twist = real(nrel_twist, c_double)
      t_nfoil = nrel_nfoil
      ! airfoil tables: 50 lines, the row count, 2 lines, rows alpha cl cd cm
      nang_max = 0
      do kf = 1, 8
         open(newunit=iu, file='./Airfoils/'//trim(foil_files(kf)), status='old', iostat=ios)
         if (ios /= 0) then
            print '(2a)', 'lbm_driver: airfoil table missing: ./Airfoils/', trim(foil_files(kf))
            stop 1
         endif
         do ia = 1, 50
            read(iu,*)
         enddo
         read(iu,*) na(kf)
         close(iu)
         nang_max = max(nang_max, na(kf))
      enddo
      allocate(t_foil(3, nang_max, 8), t_nang(8))
      t_foil = 0.0d0
      do kf = 1, 8
         open(newunit=iu, file='./Airfoils/'//trim(foil_files(kf)), status='old')
         do ia = 1, 53
            read(iu,*)
         enddo
         do ia = 1, na(kf)
            read(iu,*) row
            t_foil(:, ia, kf) = real(row(1:3), c_double)
         enddo
         close(iu)
         t_nang(kf) = na(kf)
      enddo
      th = 0.0_rk
      if (nt0 > 0) then
         write(fn,'(a,i6.6,a)') 'restart/theta_0000_', nt0, '.uf'
         inquire(file=trim(fn), exist=ex)
         if (.not. ex) then
            print '(2a)', 'lbm_driver: no restart file for theta: ', trim(fn)
            stop 1
         endif
         open(newunit=iu, file=trim(fn), form='unformatted', status='old')
         read(iu) th
         close(iu)
      endif
      t_theta = real(th, c_double)
      tdesc%nturbines = nturbines; tdesc%nblades = 3; tdesc%nchords = 17; tdesc%nfoils = 8
      tdesc%nang_max = nang_max; tdesc%reserved = 0
      tdesc%hub = c_loc(t_hub); tdesc%yaw = c_loc(t_yaw); tdesc%tilt = c_loc(t_tilt)
      tdesc%omegand = c_loc(t_omegand); tdesc%pitch = c_loc(t_pitch); tdesc%theta = c_loc(t_theta)
      tdesc%relm = c_loc(t_relm); tdesc%dc = c_loc(t_dc); tdesc%chord = c_loc(t_chord); tdesc%twist = c_loc(t_twist)
      tdesc%nfoil = c_loc(t_nfoil); tdesc%nang = c_loc(t_nang); tdesc%foil = c_loc(t_foil)
      rc = lbm_set_turbines(ctx, tdesc)
      call lbm_check(ctx, rc, 'lbm_set_turbines')
   end subroutine

   subroutine save_averages(itw)
      ! uave, vave, wave and Ti of averaging_full (src/m_averaging_full.F90:82-95), written as an
      ! unformatted record averages/averages_0000_<it>.uf instead of the Tecplot diag file
      integer, intent(in) :: itw
      integer :: iu, navg
      character(len=100) :: fn
      real(rk), allocatable, target :: ua(:,:,:), va(:,:,:), wa(:,:,:), ti(:,:,:)
      allocate(ua(0:nx+1,0:ny+1,0:nz+1), va(0:nx+1,0:ny+1,0:nz+1), wa(0:nx+1,0:ny+1,0:nz+1), ti(0:nx+1,0:ny+1,0:nz+1))
      rc = lbm_get_averages(ctx, c_loc(ua), c_loc(va), c_loc(wa), c_loc(ti), navg, realkind)
      call lbm_check(ctx, rc, 'lbm_get_averages')
      rc = lbm_set_averaging(ctx, 0_c_int)
      call lbm_check(ctx, rc, 'lbm_set_averaging')
      call execute_command_line('mkdir -p averages')
      write(fn,'(a,i6.6,a)') 'averages/averages_0000_', itw, '.uf'
      open(newunit=iu, file=trim(fn), form='unformatted', status='replace')
      write(iu) nx, ny, nz, navg, ua, va, wa, ti
      close(iu)
      print '(a,i6,a,a)', 'Averages of ', navg, ' steps written to ', trim(fn)
   end subroutine

   subroutine generate_turbulence(itw)
      ! the fields are kept on the host too, for the turbulence restart record
      integer, intent(in) :: itw
      rc = lbm_generate_turbulence(ctx, int(nrturb, c_int), real(turbulence_ampl, c_double), &
                                   int(itw, c_int64_t) + 1_c_int64_t, c_loc(uu), c_loc(vv), c_loc(ww), c_loc(rr), realkind)
      call lbm_check(ctx, rc, 'lbm_generate_turbulence')
   end subroutine

   subroutine save_restart(itw)
      ! src/m_saverestart.F90:125-136 (and the turbulence fields, :66-83)
      integer, intent(in) :: itw
      integer :: iu
      character(len=100) :: fn
      rc = lbm_get_state(ctx, c_loc(f), realkind)
      call lbm_check(ctx, rc, 'lbm_get_state')
      call execute_command_line('mkdir -p restart')
      write(fn,'(a,i6.6,a)') 'restart/restart_0000_', itw, '.uf'
      open(newunit=iu, file=trim(fn), form='unformatted', status='replace')
      write(iu) nx, ny, nz, nl, f
      close(iu)
      if (inflowturbulence) then
         write(fn,'(a,i6.6,a)') 'restart/turbulence_0000_', itw, '.uf'
         open(newunit=iu, file=trim(fn), form='unformatted', status='replace')
         write(iu) ny, nz, nrturb, uu, vv, ww, rr
         close(iu)
      endif
      if (nturbines > 0) then
         ! src/m_saverestart.F90:118-122
         rc = lbm_get_turbine_theta(ctx, t_theta)
         call lbm_check(ctx, rc, 'lbm_get_turbine_theta')
         write(fn,'(a,i6.6,a)') 'restart/theta_0000_', itw, '.uf'
         open(newunit=iu, file=trim(fn), form='unformatted', status='replace')
         write(iu) real(t_theta, rk)
         close(iu)
      endif
   end subroutine

   subroutine testing(itw)
      ! src/m_testing.F90: FP32 state with ghosts; compare when the file exists
      integer, intent(in) :: itw
      integer :: iu
      character(len=100) :: fn
      real(kind=4), allocatable :: fh(:,:,:,:)
      real(kind=4) :: fsum, fmax, diff
      integer :: l2, i2, j2, k2
      if (.not. ltesting) return
      rc = lbm_get_state(ctx, c_loc(f), realkind)
      call lbm_check(ctx, rc, 'lbm_get_state')
      call execute_command_line('mkdir -p testing')
      write(fn,'(a,i6.6,a)') 'testing/testing_0000_', itw, '.uf'
      allocate(fh(nl,0:nx+1,0:ny+1,0:nz+1))
      inquire(file=trim(fn), exist=ex)
      if (ex) then
         open(newunit=iu, file=trim(fn), form='unformatted', status='old')
         read(iu) fh
         close(iu)
         fsum = 0.0; fmax = 0.0
         do k2 = 1, nz
         do j2 = 1, ny
         do i2 = 1, nx
         do l2 = 1, nl
            diff = abs(real(f(l2,i2,j2,k2), 4) - fh(l2,i2,j2,k2))
            fsum = fsum + diff
            fmax = max(diff, fmax)
         enddo
         enddo
         enddo
         enddo
         fsum = fsum / real(nl*(nx+2)*(ny+2)*(nz+2))
         print '(a,i4,a,g12.5,a,g12.5)', 'Total misfit ', 0, ': Mean abs error=', fsum, ' max error=', fmax
      else
         fh = real(f, 4)
         open(newunit=iu, file=trim(fn), form='unformatted', status='replace')
         print '(a,a)', 'Writing new:', trim(fn)
         write(iu) fh
         close(iu)
      endif
   end subroutine

   subroutine read_inflow()
      ! uvel_shear.dat and uvel_time.dat as src/m_inflow.F90:27-86 reads them
      real(rk) :: uh(nz), z, tmp, t
      real(rk), allocatable :: td(:), ut(:), ud(:)
      integer :: kk, n, m, iu2
      uh = 1.0_rk
      inquire(file='uvel_shear.dat', exist=ex)
      if (ex) then
         open(newunit=iu2, file='uvel_shear.dat')
         do kk = 1, nz
            read(iu2,*) n, z, uh(kk)
         enddo
         close(iu2)
      endif
      uvel_shear = uh / uh(nz)
      uvel_time = uini
      udir_time = udir
      inquire(file='uvel_time.dat', exist=ex)
      if (ex) then
         open(newunit=iu2, file='uvel_time.dat')
         n = 0
         do
            read(iu2,*,iostat=ios) tmp
            if (ios /= 0) exit
            n = n + 1
         enddo
         rewind(iu2)
         allocate(td(n), ut(n), ud(n))
         do kk = 1, n
            read(iu2,*) td(kk), ut(kk), ud(kk)
            ut(kk) = ut(kk) / p_vel
         enddo
         close(iu2)
         do kk = nt0, nt1
            t = real(kk - 1) * p_time
            if (t <= td(1)) then
               uvel_time(kk) = ut(1); udir_time(kk) = ud(1)
            else if (t >= td(n)) then
               uvel_time(kk) = ut(n); udir_time(kk) = ud(n)
            else
               do m = 1, n - 1
                  if (t >= td(m) .and. t <= td(m+1)) then
                     tmp = (t - td(m)) / (td(m+1) - td(m))
                     uvel_time(kk) = ut(m) + tmp * (ut(m+1) - ut(m))
                     udir_time(kk) = ud(m) + tmp * (ud(m+1) - ud(m))
                     exit
                  endif
               enddo
            endif
         enddo
      endif
   end subroutine

end program
