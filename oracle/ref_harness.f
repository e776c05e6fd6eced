c ----------------------------------------------------------------------
c ref_harness.f -- C-callable harness around the reference's own
c routines (fft.f, init.f, fft_init.f, set_range.f, z_trans_b.f,
c z_trans_f.f, xy_trans_b.f, xy_trans_f.f, x2z_decomp.f, z2x_decomp.f,
c vel_max.f, rhs.f, spectra.f, div_check.f), compiled from
c /root/reference by oracle/Makefile into oracle/_ref/ and linked with
c the image's MPICH (one rank: MPI is initialised as a singleton).
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

c ----------------------------------------------------------------------
c Physics inputs of sam.h's /static/ block for rhs/vel_max/spectra,
c then init.f + fft_init.f + set_range.f as sam.f:44-60 runs them.
c iprm: Nx Ny Nz n_dealias i_strat nrk_max i_cfl
c rprm: xL yL zL vis Pr grav To lapse cfl0 dt0
c ----------------------------------------------------------------------
      subroutine ref_setup_phys( iprm, rprm, idims ) bind(c)
      use iso_c_binding
      include 'sam.h'
      include 'mpif.h'
      integer(c_int) iprm(7), idims(10)
      real(c_double) rprm(10)
      integer ierr_l
      logical mpi_on

      call mpi_initialized( mpi_on, ierr_l )
      if( .not. mpi_on ) call mpi_init( ierr_l )

      Nx = iprm(1)
      Ny = iprm(2)
      Nz = iprm(3)
      n_dealias = iprm(4)
      i_strat = iprm(5)
      nrk_max = iprm(6)
      i_cfl = iprm(7)
      xL = rprm(1)
      yL = rprm(2)
      zL = rprm(3)
      vis = rprm(4)
      Pr = rprm(5)
      grav = rprm(6)
      To = rprm(7)
      lapse0 = rprm(8)
      cfl0 = rprm(9)
      dt0 = rprm(10)
      iubc_z = 0
      i_les = 0
      i_prob = 0
      myid = 0
      numprocs = 1
      i_root = 0
      l_root = .false.

      call init( .false., ierr_l )
      call fft_init( )
      call set_range( Nze, Nx_min )
c     a fresh start sets lapse (initial_field.f:37)
      lapse = lapse0
      nrk = 1
      k_sq = kx_max**2 + ky_max**2 + kz_max**2
      Lm = int( sqrt(float(k_sq)) + 0.5 ) + 1

      idims(1) = kx_max
      idims(2) = ky_max
      idims(3) = kz_max
      idims(4) = Nx_min
      idims(5) = Ny_min
      idims(6) = Nz_min
      idims(7) = Nze
      idims(8) = Lu
      idims(9) = Lr
      idims(10) = Lm
      return
      end

c ----------------------------------------------------------------------
c The first substep's right-hand side as sam.f:169-243 forms it:
c z_trans_b, x2z_decomp, xy_trans_b, then rhs.f (vel_max, products,
c xy_trans_f, z2x_decomp, z_trans_f, projection, buoyancy).
c u (Nze,Ny_min,Nx_min,Lu) complex in; r (Nze,Ny_min,Nx_min,Lu+1)
c complex out (last field = pressure); st(1:17) = u_avg(4) u_var(4)
c u_max(4) energy cfl_x cfl_y cfl_z dt.
c ----------------------------------------------------------------------
      subroutine ref_rhs( u, rout, st ) bind(c)
      use iso_c_binding
      include 'sam.h'
      complex u(Nze,Ny_min,Nx_min,Lu), rout(Nze,Ny_min,Nx_min,Lu+1)
      real(c_double) st(17)
      complex, allocatable :: r(:,:,:,:), u_sav(:,:,:,:)
      real, allocatable :: uu(:,:,:,:), work(:)
      integer isy(4)

      Lw = max( 2*Nz_min*Ny_min, 4*Nze+8, (Nx+1)*Ny, 3*Lu, Nx*Ny,
     &          Lm*12 )
      allocate( r(Nze,Ny_min,Nx_min,Lr+2), uu(Nx,Ny,Nz,Lr),
     &          work(Lw), u_sav(Nz_min,Ny_min,Nx_min,3:4) )
      r = (0.0,0.0)
      isy = 0
      nrk = 1
      call z_trans_b( u, r(1,1,1,Lu+1), Lu, isy,
     &                work(1), work(1), work(Nze+1) )
      if( i_strat .eq. 1 ) then
         u_sav(1:Nz_min,1:Ny_min,1:Nx_min,3:4) =
     &      u(1:Nz_min,1:Ny_min,1:Nx_min,3:4)
      end if
      call x2z_decomp( r(1,1,1,Lu+1), uu, Lu, uu )
      call xy_trans_b( uu, uu, Lu, work )
      call rhs( uu, uu, r, u_sav, work )
      rout(:,:,:,1:Lu+1) = r(:,:,:,1:Lu+1)
      do n=1, 4
         st(n) = 0.0
         st(4+n) = 0.0
         st(8+n) = 0.0
      end do
      do n=1, Lu
         st(n) = u_avg(n)
         st(4+n) = u_var(n)
         st(8+n) = u_max(n)
      end do
      st(13) = energy
      st(14) = cfl_x
      st(15) = cfl_y
      st(16) = cfl_z
      st(17) = dt
      deallocate( r, uu, work, u_sav )
      return
      end

c vel_max.f on a physical field (Nx,Ny,Nz,Lu); st as in ref_rhs
      subroutine ref_vel_max( p, dt_in, st ) bind(c)
      use iso_c_binding
      include 'sam.h'
      real(c_double), value :: dt_in
      real p(Nx,Ny,Nz,Lu)
      real(c_double) st(17)
      real work(Lu*6)
      dt = dt_in
      call vel_max( p, work )
      st = 0.0
      do n=1, Lu
         st(n) = u_avg(n)
         st(4+n) = u_var(n)
         st(8+n) = u_max(n)
      end do
      st(13) = energy
      st(14) = cfl_x
      st(15) = cfl_y
      st(16) = cfl_z
      st(17) = dt
      return
      end

c spectra.f of (u, r): out = ek(Lm) tk(Lm) dk(Lm) vt(Lm) sample(Lm)
c energy_w div_rms div_max n_div
      subroutine ref_spectra( u, r, out ) bind(c)
      use iso_c_binding
      include 'sam.h'
      complex u(Nze,Ny_min,Nx_min,Lu), r(Nze,Ny_min,Nx_min,Lr)
      real(c_double) out(*)
      real, allocatable :: work(:)
      allocate( work(Lm*12) )
      call spectra( u, r, out(1+4*Lm), out(1), out(1+Lm),
     &              out(1+2*Lm), out(1+3*Lm), out(1+5*Lm),
     &              out(2+5*Lm), out(3+5*Lm), n_div, 0, work )
      out(4+5*Lm) = n_div
      deallocate( work )
      return
      end

c cl_energy.f and spectra.f with the Chollet-Lesieur model, as sam.f:128-132
c and spectra.f:21,39-43 run them for i_les = 1 (init.f:88-93 sets the
c model's constants): out = as ref_spectra, then E_cl, k_cl.  The setup
c of ref_setup_phys (i_les = 0) is restored afterwards.
      subroutine ref_cl_spectra( u, r, out ) bind(c)
      use iso_c_binding
      include 'sam.h'
      complex u(Nze,Ny_min,Nx_min,Lu), r(Nze,Ny_min,Nx_min,Lr)
      real(c_double) out(*)
      real, allocatable :: work(:)
      integer ierr_l
      allocate( work(Lm*12) )
      i_les = 1
      call init( .false., ierr_l )
      lapse = lapse0
      call cl_energy( u )
      call spectra( u, r, out(1+4*Lm), out(1), out(1+Lm),
     &              out(1+2*Lm), out(1+3*Lm), out(1+5*Lm),
     &              out(2+5*Lm), out(3+5*Lm), n_div, 0, work )
      out(4+5*Lm) = n_div
      out(5+5*Lm) = E_cl
      out(6+5*Lm) = k_cl
      i_les = 0
      call init( .false., ierr_l )
      lapse = lapse0
      deallocate( work )
      return
      end

c div_check.f of u: out = div_max div_rms n_div
      subroutine ref_div_check( u, out ) bind(c)
      use iso_c_binding
      include 'sam.h'
      complex u(Nze,Ny_min,Nx_min,Lu)
      real(c_double) out(3)
      call div_check( u, dmax, drms, n_div )
      out(1) = dmax
      out(2) = drms
      out(3) = n_div
      return
      end
