! ref_harness.f90 -- bind(C) entry points around the reference's own leaf
! routines, compiled from the reference sources where they lie (see
! oracle/Makefile).  TEST INFRASTRUCTURE ONLY: it lets tests/ compare the
! C oracle with the reference routines bit for bit.  The harness itself
! computes nothing; it only copies between compact C arrays and the
! reference's padded element type and calls:
!   set_points, mortar_gauss_points  (src/Grid_Routines/stac3m_geomLib.f90)
!   set_matrices, dermat, intrpmat   (src/Spectral/stac3m/matrices.f90)
!   mxm, mxv                         (src/Spectral/matmula.f90)
!   dx3dga, dy3dga, dz3dga           (src/Spectral/stac3m/FDerivLibb.f90)
!   interp                           (src/Spectral/stac3m/FInterpLib.f90)
!   Gauss_Deriv                      (src/Parabolic3D/stac3m/SolnDerivsc.f90)
!   set_geometry                     (src/Grid_Routines/set_domain_geometry.f90)
!   dtcalc                           (src/Parabolic3D/NS/NS/stac3m/dtcalc.f90)
!   rkcoefs                          (src/Independentroutines/baselib.f90)
module dsem_ref_harness
   use iso_c_binding
   use size
   use constants
   use domain_definition
   use mortar_definition
   use FE_Data_Types
   use Order_Matrices
   use physics
   use rk_coefs
   implicit none
   type(domain), allocatable, target, save :: dd
   type(mortar), allocatable, save :: mm(:)
   type(node_vector), allocatable, save :: nodes
   type(surface), save :: surfs(max_surf)
contains

   subroutine ensure_domain(p)
      integer, intent(in) :: p
      if (.not. allocated(dd)) then
         allocate(dd)
         call zero_domain(dd)
      end if
      eps = epsilon(1.0d0)
      big = huge(1.0d0)
      dd%nc = p + 2
      dd%ncg = p + 1
      dd%type = 8
      num_ordersx = 0
      num_ordersy = 0
      num_ordersz = 0
      order_mapx = 0
      order_mapy = 0
      order_mapz = 0
      call set_points(dd)
      call set_matrices(dd)
   end subroutine ensure_domain

   ! Chebyshev points and the two operators; dmx_c(j,k) = dmx(k,j), bx_c(k,j) = bx(j,k)
   subroutine ref_ops(p, cx, cxg, dmx_c, bx_c) bind(C, name='ref_ops')
      integer(c_int), value :: p
      real(c_double) :: cx(p + 2), cxg(p + 1), dmx_c(p + 2, p + 1), bx_c(p + 1, p + 2)
      integer :: j, k
      call ensure_domain(p)
      cx = dd%cx(1:p + 2, 1)
      cxg = dd%cxg(1:p + 1, 1)
      do k = 1, p + 1
         do j = 1, p + 2
            dmx_c(j, k) = dd%dmx(k, j)
         end do
      end do
      do k = 1, p + 2
         do j = 1, p + 1
            bx_c(j, k) = dd%bx(k, j)
         end do
      end do
   end subroutine ref_ops

   subroutine ref_mortar_gauss_points(n, cxg) bind(C, name='ref_mortar_gauss_points')
      integer(c_int), value :: n
      real(c_double) :: cxg(n)
      double precision :: c(nmax, 2), w(nmax, 2)
      integer :: nn(2)
      nn = n
      call mortar_gauss_points(nn, c, w)
      cxg = c(1:n, 1)
   end subroutine ref_mortar_gauss_points

   subroutine ref_rkcoefs(ark_c, brk_c, crk_c) bind(C, name='ref_rkcoefs')
      real(c_double) :: ark_c(5), brk_c(5), crk_c(5)
      call rkcoefs(4)
      ark_c = ark
      brk_c = brk
      crk_c = crk
   end subroutine ref_rkcoefs

   ! C = A*B with the reference's unrolled small-matrix kernels (column-major)
   subroutine ref_mxm(a, n1, b, n2, c, n3) bind(C, name='ref_mxm')
      integer(c_int), value :: n1, n2, n3
      real(c_double) :: a(n1, n2), b(n2, n3), c(n1, n3)
      call mxm(a, n1, b, n2, c, n3)
   end subroutine ref_mxm

   subroutine ref_interp(p, u, ugl) bind(C, name='ref_interp')
      integer(c_int), value :: p
      real(c_double) :: u(p + 1), ugl(p + 2)
      call ensure_domain(p)
      call interp(nx, dd%bx, u, p + 1, ugl, p + 2)
   end subroutine ref_interp

   ! derivative along axis (1,2,3) of a padded (nx,ny,nz) block
   subroutine ref_d3dga(p, axis, u, ud) bind(C, name='ref_d3dga')
      integer(c_int), value :: p, axis
      real(c_double) :: u(nx, ny, nz), ud(nx, ny, nz)
      call ensure_domain(p)
      select case (axis)
      case (1)
         call dx3dga(u, dd%nc, ud, dd%ncg, -0.5d0, dd%dmx)
      case (2)
         call dy3dga(u, dd%nc, ud, dd%ncg, -0.5d0, dd%dmy)
      case default
         call dz3dga(u, dd%nc, ud, dd%ncg, -0.5d0, dd%dmz)
      end select
   end subroutine ref_d3dga

   subroutine ref_gauss_deriv(p, gmetg, wlgg, wglg, wggl, wx, wy, wz) bind(C, name='ref_gauss_deriv')
      integer(c_int), value :: p
      real(c_double) :: gmetg(3, 3, nx, ny, nz)
      real(c_double) :: wlgg(nx, ny, nz), wglg(nx, ny, nz), wggl(nx, ny, nz)
      real(c_double) :: wx(nx, ny, nz), wy(nx, ny, nz), wz(nx, ny, nz)
      call ensure_domain(p)
      call Gauss_Deriv(gmetg, wlgg, wglg, wggl, wx, wy, wz, dd%ncg, dd%dmx, dd%dmy, dd%dmz)
   end subroutine ref_gauss_deriv

   ! set_geometry of one hex8 element; every face is the master of its own mortar
   subroutine ref_geometry(p, corners, xg, jacob, gmetg, gmet, nhat, norm) bind(C, name='ref_geometry')
      integer(c_int), value :: p
      real(c_double) :: corners(3, 8)
      real(c_double) :: xg(3, nx, ny, nz), jacob(nx, ny, nz)
      real(c_double) :: gmetg(3, 3, nx, ny, nz), gmet(3, 3, nx, ny, nz)
      real(c_double) :: nhat(3, nmax, nmax, 6), norm(nmax, nmax, 6)
      integer :: ic, iface
      call ensure_domain(p)
      if (.not. allocated(mm)) allocate(mm(nmp))
      if (.not. allocated(nodes)) allocate(nodes)
      do ic = 1, 8
         nodes%node(ic)%x = corners(:, ic)
         dd%node(ic) = ic
      end do
      nodes%num_nodes = 8
      do iface = 1, 6
         dd%mortar(1, iface) = iface
         dd%mortar(2, iface) = 1
         dd%which_surface(iface) = 0
         mm(iface)%lenmortar = p + 1
      end do
      dd%jacob = 0.0d0
      dd%gmet = 0.0d0
      dd%gmetg = 0.0d0
      call set_geometry(nodes, dd, mm, 0, surfs)
      xg = dd%xg
      jacob = dd%jacob
      gmetg = dd%gmetg
      gmet = dd%gmet
      do iface = 1, 6
         nhat(:, :, :, iface) = mm(iface)%n_hat
         norm(:, :, iface) = mm(iface)%norm
      end do
   end subroutine ref_geometry

   subroutine ref_dtcalc(p, gam, q, jacob, gmet, dt) bind(C, name='ref_dtcalc')
      integer(c_int), value :: p
      real(c_double), value :: gam
      real(c_double) :: q(nx, ny, nz, neq), jacob(nx, ny, nz), gmet(3, 3, nx, ny, nz)
      real(c_double) :: dt
      call ensure_domain(p)
      gamma = gam
      dd%Q = q
      dd%jacob = jacob
      dd%gmet = gmet
      call dtcalc(dt, dd)
   end subroutine ref_dtcalc

   ! gaussq (src/Independentroutines/legendre_routines.f90:98) with the
   ! arguments of the TGV statistics (BackstepProblem-stoch.f90:327):
   ! kind = 1 (Legendre), no fixed end points.
   subroutine ref_gaussq(n, t, w) bind(C, name='ref_gaussq')
      integer(c_int), value :: n
      real(c_double) :: t(n), w(n)
      real(c_double) :: b(n), endpts(2)
      endpts = 0.0d0
      call gaussq(1, n, 0.0d0, 0.0d0, 0, endpts, b, t, w)
   end subroutine ref_gaussq

   ! polyn (src/Independentroutines/baselib.f90:18-48): Lagrange basis k of
   ! the n points z at x.
   real(c_double) function ref_polyn(k, x, n, z) bind(C, name='ref_polyn')
      integer(c_int), value :: k, n
      real(c_double), value :: x
      real(c_double) :: z(n)
      double precision, external :: polyn
      ref_polyn = polyn(k, x, n, z)
   end function ref_polyn

   ! Byte sizes of the restart records write_restart writes
   ! (src/Parabolic3D/stac3m/restart.f90:59-207), in its order, from the
   ! reference's compiled declarations of TYPE(domain), TYPE(mortar),
   ! Order_Matrices and Material_Properties (src/paraFMDF/initializationsc2.f90).
   ! The harness only measures the declared components; it writes nothing.
   ! Returns the number of sizes stored in out (at most nmax_out).
   integer(c_int) function ref_restart_record_sizes(out, nmax_out) bind(C, name='ref_restart_record_sizes')
      use Material_Properties
      integer(c_int), value :: nmax_out
      integer(c_int) :: out(nmax_out)
      type(mortar), save :: mt
      integer :: n, ngrid, nmort
      double precision :: time
      n = 0
      call put(storage_size(ngrid) / 8 + storage_size(nmort) / 8)
      call put(3 * storage_size(num_ordersx) / 8)
      call put(storage_size(order_mapx) / 8 * (size(order_mapx) + size(order_mapy) + size(order_mapz)))
      call put(storage_size(bx) / 8 * size(bx))
      call put(storage_size(by) / 8 * size(by))
      call put(storage_size(bz) / 8 * size(bz))
      call put(storage_size(dmx) / 8 * size(dmx))
      call put(storage_size(dmy) / 8 * size(dmy))
      call put(storage_size(dmz) / 8 * size(dmz))
      call put(storage_size(num_mrtr_matrices) / 8)
      call put(storage_size(mrtr_mat_map) / 8 * size(mrtr_mat_map))
      call put(storage_size(prolong) / 8 * size(prolong))
      call put(storage_size(restrict) / 8 * size(restrict))
      if (.not. allocated(dd)) then
         allocate(dd)
      end if
      call put(storage_size(dd%type) / 8)
      call put(storage_size(dd%node) / 8 * size(dd%node))
      call put(storage_size(dd%orientation) / 8 * size(dd%orientation))
      call put(storage_size(dd%mortar) / 8 * size(dd%mortar))
      call put(storage_size(dd%which_surface) / 8 * size(dd%which_surface))
      call put(storage_size(dd%nc) / 8 * (size(dd%nc) + size(dd%ncg)))
      call put(storage_size(dd%Q) / 8 * size(dd%Q))
      call put(storage_size(dd%Qlgg) / 8 * size(dd%Qlgg))
      call put(storage_size(dd%Qglg) / 8 * size(dd%Qglg))
      call put(storage_size(dd%Qggl) / 8 * size(dd%Qggl))
      call put(storage_size(dd%f) / 8 * size(dd%f))
      call put(storage_size(dd%g) / 8 * size(dd%g))
      call put(storage_size(dd%h) / 8 * size(dd%h))
      call put(storage_size(dd%g_Q) / 8 * size(dd%g_Q))
      call put(storage_size(dd%material_id) / 8)
      call put(storage_size(dd%cx) / 8 * (size(dd%cx) + size(dd%cxg)))
      call put(storage_size(dd%xg) / 8 * size(dd%xg))
      call put(storage_size(dd%jacob) / 8 * size(dd%jacob))
      call put(storage_size(dd%gmet) / 8 * size(dd%gmet))
      call put(storage_size(dd%gmetg) / 8 * size(dd%gmetg))
      call put(storage_size(dd%corner) / 8 * size(dd%corner))
      call put(storage_size(dd%ibtype) / 8 * size(dd%ibtype))
      call put(storage_size(dd%bcond) / 8 * size(dd%bcond))
      call put(storage_size(dd%domain_type) / 8)
      call put(storage_size(mt%lenmortar) / 8 * size(mt%lenmortar))
      call put(storage_size(mt%len) / 8 * size(mt%len))
      call put(storage_size(mt%id) / 8 * size(mt%id))
      call put(storage_size(mt%iface) / 8 * size(mt%iface))
      call put(storage_size(mt%orient) / 8)
      call put(storage_size(mt%sign) / 8)
      call put(storage_size(mt%nsign) / 8)
      call put(storage_size(mt%n_hat) / 8 * size(mt%n_hat))
      call put(storage_size(mt%norm) / 8 * size(mt%norm))
      call put(storage_size(mt%Q) / 8 * size(mt%Q))
      call put(storage_size(mt%fv) / 8 * size(mt%fv))
      call put(storage_size(mt%xg) / 8 * size(mt%xg))
      call put(storage_size(mt%conforming) / 8 * size(mt%conforming))
      call put(storage_size(num_prop) / 8)
      call put(storage_size(material_property) / 8 * size(material_property, 1))
      call put(storage_size(time) / 8)
      ref_restart_record_sizes = n
   contains
      subroutine put(v)
         integer, intent(in) :: v
         n = n + 1
         if (n <= nmax_out) out(n) = v
      end subroutine put
   end function ref_restart_record_sizes

end module dsem_ref_harness
