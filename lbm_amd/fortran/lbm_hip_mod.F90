! Fortran binding of the lattice-update library (include/lbm_hip.h) through iso_c_binding.
!
! The reference's time loop (src/main.F90:290-329) calls postcoll, the inflow-turbulence
! forcing, forcings_apply, solids, boundarycond, drift and macrovars on device arrays it owns.
! With this module the driver creates one context per domain and calls lbm_step; state moves in
! and out in the reference's layout f(nl,0:nx+1,0:ny+1,0:nz+1).  Every function returns an
! integer status (0 = success); lbm_check stops with the library's message otherwise, the
! reference's convention for fatal paths.
module lbm_hip
   use, intrinsic :: iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: LBM_STORE_F32 = 0, LBM_STORE_F64 = 1, &
                                        LBM_STORE_F16S = 2, LBM_STORE_BF16S = 3
   integer(c_int), parameter, public :: LBM_BND_HALO = 3   ! jbnd of a domain tiled along j
   ! lbm_step_part: the cells next to the halo planes (then lbm_halo_pack), then the others
   integer(c_int), parameter, public :: LBM_PART_EDGE = 1, LBM_PART_INNER = 2

   type, bind(C), public :: lbm_desc
      integer(c_int32_t) :: nx, ny, nz, q, storage
      integer(c_int32_t) :: ibgk, ihrr, ivreman
      integer(c_int32_t) :: ibnd, jbnd, kbnd
      integer(c_int32_t) :: iturb_pos, store_tau, block_x, block_y
      real(c_double)     :: kinevisc, smagorinsky, rho0
   end type

   ! actuator-line turbines (lbm_set_turbines): sizes and the addresses of the tables
   type, bind(C), public :: lbm_turbines
      integer(c_int32_t) :: nturbines, nblades, nchords, nfoils, nang_max, reserved
      type(c_ptr) :: hub, yaw, tilt, omegand, pitch, theta, relm, dc, chord, twist, nfoil, nang, foil
   end type

   public :: lbm_set_turbines, lbm_get_turbine_theta
   public :: lbm_create, lbm_destroy, lbm_last_error, lbm_set_state, lbm_get_state, lbm_get_macros, &
             lbm_init_equilibrium, lbm_set_blanking, lbm_set_ablvisc, lbm_set_inflow, &
             lbm_set_inflow_series, lbm_set_turbulence, lbm_set_external_forcing, lbm_step, &
             lbm_synchronize, lbm_get_totals, lbm_check, lbm_error_message, lbm_generate_turbulence, &
             lbm_set_averaging, lbm_get_averages, lbm_stream, lbm_halo_bytes, lbm_halo_pack, lbm_halo_unpack, &
             lbm_step_part, lbm_set_buoyancy, lbm_get_totals_noalloc

   interface
      integer(c_int) function lbm_create(desc, ctx) bind(C, name='lbm_create')
         import :: c_int, c_ptr, lbm_desc
         type(lbm_desc), intent(in) :: desc
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function lbm_destroy(ctx) bind(C, name='lbm_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function lbm_last_error(ctx) bind(C, name='lbm_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function lbm_set_state(ctx, f, realkind) bind(C, name='lbm_set_state')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         type(c_ptr), value :: f
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_get_state(ctx, f, realkind) bind(C, name='lbm_get_state')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         type(c_ptr), value :: f
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_get_macros(ctx, rho, u, v, w, realkind) bind(C, name='lbm_get_macros')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, rho, u, v, w
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_init_equilibrium(ctx, rho, u, v, w, realkind, it0) bind(C, name='lbm_init_equilibrium')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, rho, u, v, w
         integer(c_int), value :: realkind, it0
      end function
      integer(c_int) function lbm_set_blanking(ctx, mask) bind(C, name='lbm_set_blanking')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, mask
      end function
      integer(c_int) function lbm_set_ablvisc(ctx, abl, realkind) bind(C, name='lbm_set_ablvisc')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, abl
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_set_inflow(ctx, shear, uini, udir, realkind) bind(C, name='lbm_set_inflow')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx, shear
         real(c_double), value :: uini, udir
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_set_inflow_series(ctx, it_lo, it_hi, ut, ud, realkind) bind(C, name='lbm_set_inflow_series')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: it_lo, it_hi
         type(c_ptr), value :: ut, ud
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_set_turbulence(ctx, uu, vv, ww, nrturb, ampl, realkind) bind(C, name='lbm_set_turbulence')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx, uu, vv, ww
         integer(c_int), value :: nrturb
         real(c_double), value :: ampl
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_generate_turbulence(ctx, nrturb, ampl, seed, uu, vv, ww, rr, realkind) &
            bind(C, name='lbm_generate_turbulence')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int), value :: nrturb
         real(c_double), value :: ampl
         integer(c_int64_t), value :: seed
         type(c_ptr), value :: uu, vv, ww, rr
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_set_external_forcing(ctx, force, realkind) bind(C, name='lbm_set_external_forcing')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, force
         integer(c_int), value :: realkind
      end function
      ! buoyancy_forcing (src/m_buoyancy_forcing.F90:3-46) from the caller's pottemp(0:nx+1,0:ny+1,0:nz+1)
      integer(c_int) function lbm_set_buoyancy(ctx, pottemp, theta0, scaling, realkind) bind(C, name='lbm_set_buoyancy')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx, pottemp
         real(c_double), value :: theta0, scaling
         integer(c_int), value :: realkind
      end function
      ! turbine_forcing (src/m_turbine_forcing.F90:8-140) on the device, every step
      integer(c_int) function lbm_set_turbines(ctx, t) bind(C, name='lbm_set_turbines')
         import :: c_int, c_ptr, lbm_turbines
         type(c_ptr), value :: ctx
         type(lbm_turbines), intent(in) :: t
      end function
      integer(c_int) function lbm_get_turbine_theta(ctx, theta) bind(C, name='lbm_get_turbine_theta')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: theta(*)
      end function
      integer(c_int) function lbm_step(ctx, it0, nsteps) bind(C, name='lbm_step')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: it0, nsteps
      end function
      ! halo planes of a domain tiled along j (jbnd = LBM_BND_HALO): after each one-step lbm_step,
      ! pack (side 0: plane 1, side 1: plane ny, 2: both) into device memory, send/receive it
      ! (GPU-aware MPI on the device buffers, or RCCL on lbm_stream), unpack into the ghost planes;
      ! replaces mpi_halo_exchange_j (src/m_mpi_halo_exchange_j.F90:55-120)
      type(c_ptr) function lbm_stream(ctx) bind(C, name='lbm_stream')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function lbm_halo_bytes(ctx, nbytes) bind(C, name='lbm_halo_bytes')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: ctx
         integer(c_size_t), intent(out) :: nbytes
      end function
      integer(c_int) function lbm_halo_pack(ctx, side, dev_buf) bind(C, name='lbm_halo_pack')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, dev_buf
         integer(c_int), value :: side
      end function
      integer(c_int) function lbm_halo_unpack(ctx, side, dev_buf) bind(C, name='lbm_halo_unpack')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, dev_buf
         integer(c_int), value :: side
      end function
      integer(c_int) function lbm_step_part(ctx, it, part) bind(C, name='lbm_step_part')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: it, part
      end function
      integer(c_int) function lbm_set_averaging(ctx, on) bind(C, name='lbm_set_averaging')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
      end function
      integer(c_int) function lbm_get_averages(ctx, uave, vave, wave, ti, count, realkind) bind(C, name='lbm_get_averages')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, uave, vave, wave, ti
         integer(c_int), intent(out) :: count
         integer(c_int), value :: realkind
      end function
      integer(c_int) function lbm_synchronize(ctx) bind(C, name='lbm_synchronize')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function lbm_get_totals(ctx, mass, mom) bind(C, name='lbm_get_totals')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: mass
         real(c_double), intent(out) :: mom(3)
      end function
      ! lbm_get_totals without any device allocation (no counterpart in the reference)
      integer(c_int) function lbm_get_totals_noalloc(ctx, mass, mom) bind(C, name='lbm_get_totals_noalloc')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: mass
         real(c_double), intent(out) :: mom(3)
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   ! The library's message for ctx (c_null_ptr: the last failed lbm_create).
   function lbm_error_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = lbm_last_error(ctx)
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      enddo
   end function

   subroutine lbm_check(ctx, rc, what)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= 0) then
         print '(a,a,i3,a,a)', what, ' failed, status', rc, ': ', lbm_error_message(ctx)
         stop 1
      endif
   end subroutine

end module
