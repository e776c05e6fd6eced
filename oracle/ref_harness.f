c ----------------------------------------------------------------------
c ref_harness.f -- C-callable harness around the reference's own
c MPI-free routines (fft.f, init.f, fft_init.f, set_range.f,
c z_trans_b.f, z_trans_f.f, xy_trans_b.f, xy_trans_f.f), compiled
c from /root/reference by oracle/Makefile into oracle/_ref/.
c
c TEST INFRASTRUCTURE ONLY: used by tests/ to pin the C oracle.  The
c harness only fills the reference's COMMON blocks (sam.h) for one
c rank and forwards array arguments; all arithmetic is the
c reference's.
c ----------------------------------------------------------------------

      subroutine ref_setup( nx_in, ny_in, nz_in, xl_in, yl_in, zl_in,
     &                      ndeal, istrat, idims ) bind(c)
      use iso_c_binding
      include 'sam.h'
      integer(c_int), value :: nx_in, ny_in, nz_in, ndeal, istrat
      real(c_double), value :: xl_in, yl_in, zl_in
      integer(c_int) idims(10)
      integer ierr_l

      Nx = nx_in
      Ny = ny_in
      Nz = nz_in
      xL = xl_in
      yL = yl_in
      zL = zl_in
      n_dealias = ndeal
      i_strat = istrat
      iubc_z = 0
      nrk_max = 3
      i_les = 0
      i_prob = 0
      To = 1.0
      Pr = 1.0
      grav = 0.0
      lapse0 = 1.0
      dt0 = 0.01
      myid = 0
      numprocs = 1
      i_root = 0
      l_root = .false.

      call init( .false., ierr_l )
      call fft_init( )
      call set_range( Nze, Nx_min )

      idims(1) = kx_max
      idims(2) = ky_max
      idims(3) = kz_max
      idims(4) = Nx_min
      idims(5) = Ny_min
      idims(6) = Nz_min
      idims(7) = Nze
      idims(8) = Lu
      idims(9) = Lr
      idims(10) = ierr_l
      return
      end

      subroutine ref_z_trans_b( u1, u2, nvar ) bind(c)
      use iso_c_binding
      include 'sam.h'
      integer(c_int), value :: nvar
      real(c_double) u1(*), u2(*)
      integer isy(16)
      real, allocatable :: wk(:)
      allocate( wk(4*Nze+8) )
      isy = 0
      call z_trans_b( u1, u2, nvar, isy, wk(1), wk(1), wk(2*Nze+3) )
      deallocate( wk )
      return
      end

      subroutine ref_xy_trans_b( u1, u2, nvar ) bind(c)
      use iso_c_binding
      include 'sam.h'
      integer(c_int), value :: nvar
      real(c_double) u1(*), u2(*)
      real, allocatable :: wk(:)
      allocate( wk((Nx+1)*Ny) )
      call xy_trans_b( u1, u2, nvar, wk )
      deallocate( wk )
      return
      end

      subroutine ref_xy_trans_f( u1, u2, nvar ) bind(c)
      use iso_c_binding
      include 'sam.h'
      integer(c_int), value :: nvar
      real(c_double) u1(*), u2(*)
      real, allocatable :: wk(:)
      allocate( wk(2*Nx*Ny+16) )
      call xy_trans_f( u1, u2, nvar, wk )
      deallocate( wk )
      return
      end

      subroutine ref_z_trans_f( u, nvar ) bind(c)
      use iso_c_binding
      include 'sam.h'
      integer(c_int), value :: nvar
      real(c_double) u(*)
      integer isy(16)
      real, allocatable :: wk(:)
      allocate( wk(2*Nze+8) )
      isy = 0
      call z_trans_f( u, nvar, isy, wk(1), wk(Nze+1) )
      deallocate( wk )
      return
      end

c     FFTPACK entry points.  dir < 0: forward, dir > 0: backward.
      subroutine ref_rfft( n, r, idir ) bind(c)
      use iso_c_binding
      integer(c_int), value :: n, idir
      real(c_double) r(*)
      real(8), allocatable :: ws(:)
      allocate( ws(2*n+15) )
      call rffti( n, ws )
      if( idir .lt. 0 ) then
         call rfftf( n, r, ws )
      else
         call rfftb( n, r, ws )
      end if
      deallocate( ws )
      return
      end

      subroutine ref_cfft( n, c, idir ) bind(c)
      use iso_c_binding
      integer(c_int), value :: n, idir
      real(c_double) c(*)
      real(8), allocatable :: ws(:)
      allocate( ws(4*n+15) )
      call cffti( n, ws )
      if( idir .lt. 0 ) then
         call cfftf( n, c, ws )
      else
         call cfftb( n, c, ws )
      end if
      deallocate( ws )
      return
      end
