! samhip_mod.f90 -- iso_c_binding interface to libsamhip (include/samhip.h).
!
! The derived types mirror the C structs field for field; every entry point
! returns the library's int status (0 = ok).
module samhip_mod
  use iso_c_binding
  implicit none

  integer, parameter :: SAM_OK = 0, SAM_ERR_REQUIRED = 6, SAM_ERR_VALUE = 9, &
       SAM_ERR_SIZE = 10, SAM_ERR_UNSUPPORTED = 20
  integer, parameter :: SAM_PREC_FP64 = 0, SAM_PREC_SPLIT = 1
  integer, parameter :: SAM_SGS_NONE = 0, SAM_SGS_SMAG = 1, SAM_SGS_DYN = 2, SAM_SGS_MLP = 3
  integer, parameter :: SAM_LABELS_FILTERED = 0, SAM_LABELS_DYNAMIC = 1, SAM_MLP_NPARAM = 10448

  type, bind(c) :: sam_params
     integer(c_int) :: nx, ny, nz, n_dealias, i_strat, nrk_max, i_cfl, iubc_z, precision, sgs
     real(c_double) :: xl, yl, zl, vis, pr, grav, to, lapse, cfl0, dt0, c_smag, pr_t
  end type sam_params

  type, bind(c) :: sam_dims
     integer(c_int) :: kx_max, ky_max, kz_max, nx_min, ny_min, nz_min, nze, lu, lr, lm
  end type sam_dims

  type, bind(c) :: sam_diag
     integer(c_long_long) :: nt
     real(c_double) :: time, dt
     real(c_double) :: u_avg(4), u_var(4), u_max(4)
     real(c_double) :: energy, cfl_x, cfl_y, cfl_z, c_delta_sq
     integer(c_int) :: sgs_active, pad_
  end type sam_diag

  interface
     integer(c_int) function sam_create(p, h) bind(c, name='sam_create')
       import :: c_int, c_ptr, sam_params
       type(sam_params), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function sam_create

     integer(c_int) function sam_destroy(h) bind(c, name='sam_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function sam_destroy

     integer(c_int) function sam_get_dims(h, d) bind(c, name='sam_get_dims')
       import :: c_int, c_ptr, sam_dims
       type(c_ptr), value :: h
       type(sam_dims), intent(out) :: d
     end function sam_get_dims

     type(c_ptr) function sam_last_error() bind(c, name='sam_last_error')
       import :: c_ptr
     end function sam_last_error

     integer(c_int) function sam_set_state(h, u, ldz) bind(c, name='sam_set_state')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: u(*)
       integer(c_int), value :: ldz
     end function sam_set_state

     integer(c_int) function sam_get_state(h, u, ldz) bind(c, name='sam_get_state')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(out) :: u(*)
       integer(c_int), value :: ldz
     end function sam_get_state

     ! the velocity SAM writes inside a step (after substep nrk's zeta update)
     integer(c_int) function sam_get_state_substep(h, nrk, u, ldz) bind(c, name='sam_get_state_substep')
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nrk, ldz
       complex(c_double_complex), intent(out) :: u(*)
     end function sam_get_state_substep

     ! stage entry points in the reference's layouts (include/samhip.h)
     integer(c_int) function sam_z_trans_b(h, u, ldz, out, nvar) bind(c, name='sam_z_trans_b')
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: ldz, nvar
       complex(c_double_complex), intent(in) :: u(*)
       complex(c_double_complex), intent(out) :: out(*)
     end function sam_z_trans_b

     integer(c_int) function sam_x2z(h, in, out, nvar) bind(c, name='sam_x2z')
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nvar
       complex(c_double_complex), intent(in) :: in(*)
       complex(c_double_complex), intent(out) :: out(*)
     end function sam_x2z

     integer(c_int) function sam_xy_trans_b(h, in, out, nvar) bind(c, name='sam_xy_trans_b')
       import :: c_ptr, c_int, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nvar
       complex(c_double_complex), intent(in) :: in(*)
       real(c_double), intent(out) :: out(*)
     end function sam_xy_trans_b

     integer(c_int) function sam_xy_trans_f(h, in, out, nvar) bind(c, name='sam_xy_trans_f')
       import :: c_ptr, c_int, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nvar
       real(c_double), intent(in) :: in(*)
       complex(c_double_complex), intent(out) :: out(*)
     end function sam_xy_trans_f

     integer(c_int) function sam_z2x(h, in, out, nvar) bind(c, name='sam_z2x')
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nvar
       complex(c_double_complex), intent(in) :: in(*)
       complex(c_double_complex), intent(out) :: out(*)
     end function sam_z2x

     integer(c_int) function sam_z_trans_f(h, in, out, nvar) bind(c, name='sam_z_trans_f')
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nvar
       complex(c_double_complex), intent(in) :: in(*)
       complex(c_double_complex), intent(out) :: out(*)
     end function sam_z_trans_f

     ! r_prev / r may be c_null_ptr: pass c_loc of the arrays otherwise
     integer(c_int) function sam_rk_update(h, nrk, dt, u, r_prev, r, ldz) bind(c, name='sam_rk_update')
       import :: c_ptr, c_int, c_double, c_double_complex
       type(c_ptr), value :: h, r_prev, r
       integer(c_int), value :: nrk, ldz
       real(c_double), value :: dt
       complex(c_double_complex), intent(inout) :: u(*)
     end function sam_rk_update

     integer(c_int) function sam_div_check(h, out) bind(c, name='sam_div_check')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(3)
     end function sam_div_check

     integer(c_int) function sam_set_clock(h, time, dt, nt) bind(c, name='sam_set_clock')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: h
       real(c_double), value :: time, dt
       integer(c_long_long), value :: nt
     end function sam_set_clock

     integer(c_int) function sam_step(h, nsteps, want_spectra, diag) bind(c, name='sam_step')
       import :: c_int, c_ptr, sam_diag
       type(c_ptr), value :: h
       integer(c_int), value :: nsteps, want_spectra
       type(sam_diag), intent(out) :: diag
     end function sam_step

     integer(c_int) function sam_rhs(h, r, ldz, want_spectra, diag) bind(c, name='sam_rhs')
       import :: c_int, c_ptr, sam_diag
       type(c_ptr), value :: h
       type(c_ptr), value :: r
       integer(c_int), value :: ldz, want_spectra
       type(sam_diag), intent(out) :: diag
     end function sam_rhs

     integer(c_int) function sam_get_spectra(h, out) bind(c, name='sam_get_spectra')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(*)
     end function sam_get_spectra

     ! learned closure (no counterpart in the reference)
     integer(c_int) function sam_mlp_create(m, seed, cap) bind(c, name='sam_mlp_create')
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), intent(out) :: m
       integer(c_int), value :: seed
       integer(c_long_long), value :: cap
     end function sam_mlp_create

     integer(c_int) function sam_mlp_destroy(m) bind(c, name='sam_mlp_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: m
     end function sam_mlp_destroy

     integer(c_int) function sam_mlp_set_params(m, params) bind(c, name='sam_mlp_set_params')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: m
       real(c_float), intent(in) :: params(*)
     end function sam_mlp_set_params

     integer(c_int) function sam_mlp_set_labels(m, mode, pr_t) bind(c, name='sam_mlp_set_labels')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: m
       integer(c_int), value :: mode
       real(c_double), value :: pr_t
     end function sam_mlp_set_labels

     integer(c_int) function sam_mlp_set_activation(m, act) bind(c, name='sam_mlp_set_activation')
       import :: c_int, c_ptr
       type(c_ptr), value :: m
       integer(c_int), value :: act
     end function sam_mlp_set_activation

     integer(c_int) function sam_attach_mlp(h, m) bind(c, name='sam_attach_mlp')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, m
     end function sam_attach_mlp

     ! per-step extras: force.f and the Chollet-Lesieur model
     integer(c_int) function sam_set_forcing(h, k_force) bind(c, name='sam_set_forcing')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: k_force
     end function sam_set_forcing

     integer(c_int) function sam_set_cl(h, mode) bind(c, name='sam_set_cl')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: mode
     end function sam_set_cl

     ! planes output (write_planes.f)
     integer(c_int) function sam_set_planes(h, keep) bind(c, name='sam_set_planes')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: keep
     end function sam_set_planes

     integer(c_int) function sam_planes(h, n_xy, k_xy, n_xz, j_xz, n_yz, i_yz, mean, xy, xz, yz) &
          bind(c, name='sam_planes')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: n_xy, n_xz, n_yz
       integer(c_int), intent(in) :: k_xy(*), j_xz(*), i_yz(*)
       real(c_double), intent(out) :: mean(*), xy(*), xz(*), yz(*)
     end function sam_planes

     integer(c_int) function sam_set_mlp_refresh(h, nsteps) bind(c, name='sam_set_mlp_refresh')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nsteps
     end function sam_set_mlp_refresh
  end interface

contains

  ! message of the library's last error as a Fortran string
  function sam_error_message() result(msg)
    character(len=512) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    p = sam_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function sam_error_message

end module samhip_mod
