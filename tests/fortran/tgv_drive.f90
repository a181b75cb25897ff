! Fortran driver of the GPU library through the iso_c_binding module: reads
! a mesh and a J*Q state in the reference's padded layout, takes nsteps
! steps with the tstep-compatible wrapper, takes the TGV statistics after
! every step (turbk, enstrophy dissipation) and writes the compact state.
! Input/output are stream files written/read by tests/test_gpu_fortran.py.
program tgv_drive
   use iso_c_binding
   use dsem_hip
   implicit none
   integer(c_int32_t) :: hdr(4)
   integer(c_int32_t), allocatable, target :: me(:), mf(:), ms(:), mn(:)
   real(c_double), allocatable, target :: corners(:), jq_pad(:), jq_out(:)
   real(c_double) :: phys(4), dt, resid, t, turbk, eps
   real(c_double), allocatable :: dtv(:)
   integer(c_int) :: rc, ierr, nsteps, k
   type(dsem_mesh_desc) :: desc
   type(dsem_params) :: prm
   type(c_ptr) :: ctx
   character(len=512) :: fin, fout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr                          ! K, p, nmortar, nsteps
   nsteps = hdr(4)
   allocate(me(2*hdr(3)), mf(2*hdr(3)), ms(hdr(3)), mn(hdr(3)))
   allocate(corners(24*hdr(1)), jq_pad(hdr(1)*5*1000), jq_out(hdr(1)*5*(hdr(2)+1)**3))
   read(10) me, mf, ms, mn, corners, jq_pad, phys, dt
   close(10)

   desc%K = hdr(1); desc%p = hdr(2); desc%nmortar = hdr(3); desc%reserved = 0
   desc%mortar_elem = c_loc(me); desc%mortar_face = c_loc(mf)
   desc%mortar_sign = c_loc(ms); desc%mortar_nsign = c_loc(mn)
   desc%mortar_orient = c_null_ptr; desc%corners = c_loc(corners)
   prm%gamma = phys(1); prm%re = phys(2); prm%pr = phys(3); prm%mach = phys(4)
   prm%shock = 0; prm%precision = DSEM_PREC_FP64; prm%metrics = DSEM_METRICS_PER_POINT; prm%ev_flags = 0
   prm%cmu = 1.0d0; prm%mumax = 1.0d0; prm%cmax = 0.5d0

   rc = dsem_create(desc, prm, 0, ctx)
   if (rc /= DSEM_OK) then
      print *, 'dsem_create failed: ', trim(dsem_last_error(ctx))
      stop 1
   end if
   rc = dsem_set_state(ctx, jq_pad, 1)
   if (rc /= DSEM_OK) stop 2
   rc = dsem_keep_lagged_states(ctx, 1)
   if (rc /= DSEM_OK) stop 5
   allocate(dtv(hdr(1)))
   dtv = dt
   t = 0.0d0
   do k = 1, nsteps
      call dsem_tstep(ctx, t, dtv, resid, ierr)
      if (ierr /= DSEM_OK) then
         print *, 'tstep failed: ', trim(dsem_last_error(ctx))
         stop 3
      end if
      call dsem_energy(ctx, turbk, eps, ierr)
      if (ierr /= DSEM_OK) then
         print *, 'energy failed: ', trim(dsem_last_error(ctx))
         stop 6
      end if
      t = t + dt
   end do
   rc = dsem_get_state(ctx, jq_out, 0)
   if (rc /= DSEM_OK) stop 4
   rc = dsem_destroy(ctx)
   open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(11) resid, turbk, eps, jq_out
   close(11)
   print '(a,i0,a,es14.6,a,es22.15)', 'fortran driver: ', nsteps, ' steps, resid ', resid, ' turbk ', turbk
end program tgv_drive
