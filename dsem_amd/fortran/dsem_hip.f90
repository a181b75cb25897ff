! dsem_hip.f90 -- Fortran iso_c_binding interface to libdsem_hip.so, the
! MI355X DSEM right-hand side and RK4(5) step (C ABI in include/dsem_hip.h).
!
! A driver that calls the reference's
!     CALL tstep(time, delta_t, d, ngridl, mrtr, nmort, resid)
! (src/3DDriver_parfmdf.f90:466, src/tstep.f90:12) keeps its calling
! convention through dsem_tstep below: the state lives on the GPU between
! calls, and is gathered/scattered with dsem_set_state / dsem_get_state
! (layout 1 is the reference's d(id)%Q(nx,ny,nz,neq) block with nx=10)
! only at start-up and output time.
module dsem_hip
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: DSEM_OK = 0
   integer(c_int), parameter :: DSEM_ERR_ARG = -1
   integer(c_int), parameter :: DSEM_ERR_HIP = -2
   integer(c_int), parameter :: DSEM_ERR_NONFINITE = -3
   integer(c_int), parameter :: DSEM_ERR_UNSUPPORTED = -4
   integer(c_int), parameter :: DSEM_ERR_NOMEM = -5
   integer(c_int), parameter :: DSEM_PREC_FP64 = 0, DSEM_PREC_FP32 = 1, DSEM_PREC_BF16X3 = 2
   integer(c_int), parameter :: DSEM_METRICS_PER_POINT = 0, DSEM_METRICS_AUTO = 1, DSEM_METRICS_PER_ELEMENT = 2
   integer(c_int), parameter :: DSEM_EV_DUCROS = 1, DSEM_EV_SMOOTHING = 2, DSEM_EV_CMAX_CAP = 4

   type, bind(C) :: dsem_mesh_desc
      integer(c_int32_t) :: K, p, nmortar, reserved
      type(c_ptr) :: mortar_elem, mortar_face, mortar_sign, mortar_nsign, mortar_orient
      type(c_ptr) :: corners
   end type dsem_mesh_desc

   type, bind(C) :: dsem_params
      real(c_double) :: gamma, re, pr, mach
      integer(c_int32_t) :: shock, precision, metrics, ev_flags
      real(c_double) :: cmu, mumax, cmax
   end type dsem_params

   interface
      integer(c_int) function dsem_create(mesh, params, device, ctx) bind(C, name='dsem_create')
         import :: c_int, c_ptr, dsem_mesh_desc, dsem_params
         type(dsem_mesh_desc), intent(in) :: mesh
         type(dsem_params), intent(in) :: params
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: ctx
      end function dsem_create

      integer(c_int) function dsem_destroy(ctx) bind(C, name='dsem_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function dsem_destroy

      type(c_ptr) function dsem_last_error_c(ctx) bind(C, name='dsem_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function dsem_last_error_c

      integer(c_int) function dsem_set_state(ctx, JQ, layout) bind(C, name='dsem_set_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: JQ(*)
         integer(c_int), value :: layout
      end function dsem_set_state

      integer(c_int) function dsem_get_state(ctx, JQ, layout) bind(C, name='dsem_get_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: JQ(*)
         integer(c_int), value :: layout
      end function dsem_get_state

      integer(c_int) function dsem_rhs(ctx, t, Qt) bind(C, name='dsem_rhs')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: t
         real(c_double), intent(out) :: Qt(*)
      end function dsem_rhs

      integer(c_int) function dsem_rk_step(ctx, t, dt, resid) bind(C, name='dsem_rk_step')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: t, dt
         real(c_double), intent(out) :: resid
      end function dsem_rk_step

      integer(c_int) function dsem_dtcalc(ctx, dt) bind(C, name='dsem_dtcalc')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: dt
      end function dsem_dtcalc

      integer(c_int) function dsem_gauss_points(ctx, xg, jacob) bind(C, name='dsem_gauss_points')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: xg(*), jacob(*)
      end function dsem_gauss_points

      integer(c_int) function dsem_set_viscosity(ctx, muh) bind(C, name='dsem_set_viscosity')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: muh(*)
      end function dsem_set_viscosity

      ! turbk of do_user_stuff and eps of energydissipation (either may be
      ! c_null_ptr through the c_loc of an absent target; see dsem_energy)
      integer(c_int) function dsem_reduce_ke(ctx, Ek, eps) bind(C, name='dsem_reduce_ke')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, Ek, eps
      end function dsem_reduce_ke

      ! static Smagorinsky model (input keyword 'smagorinsky', physics C_s, numax)
      integer(c_int) function dsem_set_les(ctx, enable, cs, numax) bind(C, name='dsem_set_les')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable
         real(c_double), value :: cs, numax
      end function dsem_set_les

      integer(c_int) function dsem_keep_lagged_states(ctx, enable) bind(C, name='dsem_keep_lagged_states')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable
      end function dsem_keep_lagged_states

      ! Smagorinsky sensor switches (input keywords shockSensor, rhoSensor):
      ! DSEM_LES_SHOCK_SENSOR = 1, DSEM_LES_RHO_SENSOR = 2
      integer(c_int) function dsem_set_les_sensors(ctx, sensors) bind(C, name='dsem_set_les_sensors')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: sensors
      end function dsem_set_les_sensors

      ! modalfilter(d,ngridl,time,k), called by the driver when mod(k,50) == 0
      ! (src/3DDriver_parfmdf.f90:462-463); cutoff = p+1 keeps every mode
      integer(c_int) function dsem_modal_filter(ctx, cutoff) bind(C, name='dsem_modal_filter')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: cutoff
      end function dsem_modal_filter
   end interface

contains

   ! Message of the last failed call on ctx.
   function dsem_last_error(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=256) :: msg
      character(kind=c_char), pointer :: s(:)
      integer :: i
      msg = ' '
      call c_f_pointer(dsem_last_error_c(ctx), s, [256])
      do i = 1, 256
         if (s(i) == c_null_char) exit
         msg(i:i) = s(i)
      end do
   end function dsem_last_error

   ! Drop-in for tstep(time,dt,d,ngrid,mrtr,nmort,resid): one RK4(5) step of
   ! the state held by ctx.  dt is the reference's per-element array with the
   ! global step in every entry (time-accurate mode); a NaN/Inf residual is
   ! reported through ierr = DSEM_ERR_NONFINITE.
   subroutine dsem_tstep(ctx, time, dt, resid, ierr)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in) :: time
      real(c_double), intent(in) :: dt(:)
      real(c_double), intent(out) :: resid
      integer(c_int), intent(out) :: ierr
      ierr = dsem_rk_step(ctx, time, dt(1), resid)
   end subroutine dsem_tstep

   ! Statistics of the TGV driver after a step: turbk (fort.22) and the
   ! enstrophy dissipation of energydissipation (the `dissipated` file);
   ! the caller forms dissp = (prestep - turbk)/dt as do_user_stuff does.
   ! eps needs dsem_keep_lagged_states(ctx, 1) before the step.
   subroutine dsem_energy(ctx, turbk, eps, ierr)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(out), target :: turbk, eps
      integer(c_int), intent(out) :: ierr
      ierr = dsem_reduce_ke(ctx, c_loc(turbk), c_loc(eps))
   end subroutine dsem_energy

end module dsem_hip
