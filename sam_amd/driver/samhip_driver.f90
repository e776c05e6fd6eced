! samhip_driver.f90 -- SAM's run driver on top of libsamhip.
!
! Reads input.dat in SAM's label/value format (read_line.f:24-89,
! set_labels.f:29-102, check_inputs.f:14-231), restarts from header.NNNNNN /
! vel.NNNNNN (read_header.f, read_field.f), advances the RK3 loop of
! sam.f:122-326 on the GPU through the C ABI, and writes energy.out,
! umax.out, tstep.out, variance.out, spectra.out (sam.f:248-273,
! spectra.f:92-111) and vel.NNNNNN / header.NNNNNN (write_field.f,
! write_header.f) in the reference's formats.
!
! Usage:  samhip_driver [input file, default input.dat]
!
! Labels beyond the reference's (warned about as extraneous by SAM itself):
!   i_prec  0 FP64 (default), 1 FP32 transforms with FP64 state
!   i_lapse 0 restart keeps lapse = 0 as SAM does (initial_field.f:37 is the
!           only place lapse is set), 1 uses lapse0 after a restart
!   pr_t    turbulent Prandtl number of the SGS scalar flux (default 1)
!   mlp_labels   i_les 4: 1 (default) the network was trained on unit-
!                coefficient dynamic-model labels and is scaled by C Delta^2
!                from the dynamic procedure; 0 trained on box-filter labels
!   mlp_refresh  i_les 4, mlp_labels 1: steps between C Delta^2 updates (20)
!   mlp_act      i_les 4: hidden activation the network was trained with,
!                0 ReLU (default), 1 GELU (tanh form)
!   cl_couple    i_les 1: 0 (default) the Chollet-Lesieur nu_t(k) enters the
!                spectra only, as in SAM; 1 also the integrating factor
!   vel_midstep  vel.NNNNNN written for nt < nt_end: 0 (default) the state at
!                the start of step nt; 1 what SAM writes there, the velocity
!                of the step's last substep after its zeta update
!                (sam.f:210-221,277-284)
! n_skip_p: planes every n_skip_p steps (write_planes.f): time.out, mean.out,
!           xy[m].dat, xz[m].dat, yz[m].dat at k_xy_plane[m], j_xz_plane[m],
!           i_yz_plane[m] (first plane of each kind N/2 when not given)
! i_force 1: force.f after every step (shells 2..k_force+1 held at the last
!            spectrum written to spectra.out)
! i_les: 0 DNS, 1 Chollet-Lesieur (diagnostics of spectra.f; the step is DNS),
!        2 Smagorinsky (sm.f), 3 dynamic Smagorinsky (dm.f), 4 learned closure
!        with its parameters read from mlp.params (SAM_MLP_NPARAM + 32 raw
!        FP32 values as written by sam_mlp_get_params).
program samhip_driver
  use iso_c_binding
  use samhip_mod
  implicit none

  integer, parameter :: NIN = 72, NDYN = 6, NPAR = NIN + NDYN, NEXT = 8
  character(len=12) :: labels(NPAR + NEXT)
  real(8) :: values(NPAR + NEXT), defaults(NPAR + NEXT)
  logical :: required(NPAR + NEXT), fixed(NPAR + NEXT), found(NPAR + NEXT)
  real(8) :: hvalues(NPAR + NEXT)
  logical :: hfound(NPAR + NEXT)

  character(len=256) :: fin
  character(len=6) :: ext
  integer :: ierr, nt, nt_start, nt_end, n, i, lu, lm, want
  integer(c_int) :: rc
  type(sam_params) :: p
  type(sam_dims) :: d
  type(sam_diag) :: dg
  type(c_ptr) :: h, mlp
  real(c_float), allocatable :: mlp_params(:)
  complex(c_double_complex), allocatable :: u(:, :, :, :)
  real(c_double), allocatable :: spec(:)
  real(8) :: time, dt, t_stat
  integer :: n_skip_h, n_skip_v, n_steps, i_restart, nt_restart, i_les, i_strat
  logical :: appendf
  ! planes output (write_planes.f)
  logical :: planes_on
  integer :: n_skip_p, n_frame_p, n_xy, n_xz, n_yz
  integer(c_int) :: k_xy(9), j_xz(9), i_yz(9)
  real(c_double), allocatable :: pmean(:, :), pxy(:, :, :, :), pxz(:, :, :, :), pyz(:, :, :, :)

  fin = 'input.dat'
  if (command_argument_count() >= 1) call get_command_argument(1, fin)

  call set_labels()
  values = defaults
  call read_params(trim(fin), values, found, ierr)
  if (ierr /= 0) call fail(ierr, 'reading '//trim(fin))
  call check_inputs(ierr)
  if (ierr /= 0) call fail(ierr, 'input check')

  i_restart = iv('i_restart')
  nt_restart = iv('nt_restart')
  i_les = iv('i_les')
  i_strat = iv('i_strat')
  n_steps = iv('n_steps')
  n_skip_h = iv('n_skip_h')
  n_skip_v = iv('n_skip_v')

  ! ---- parameters for the library ------------------------------------
  p%nx = iv('nx'); p%ny = iv('ny'); p%nz = iv('nz')
  p%n_dealias = iv('n_dealias'); p%i_strat = i_strat; p%nrk_max = iv('nrk_max')
  p%i_cfl = iv('i_cfl'); p%iubc_z = iv('iubc_z'); p%precision = iv('i_prec')
  select case (i_les)
  case (2); p%sgs = SAM_SGS_SMAG
  case (3); p%sgs = SAM_SGS_DYN
  case (4); p%sgs = SAM_SGS_MLP
  case default; p%sgs = SAM_SGS_NONE
  end select
  p%xl = rv('xl'); p%yl = rv('yl'); p%zl = rv('zl')
  p%vis = rv('vis'); p%pr = rv('pr'); p%grav = rv('grav'); p%to = rv('to')
  p%lapse = rv('lapse0')
  if (i_restart == 1 .and. iv('i_lapse') == 0 .and. i_strat == 1) then
     print *, 'WARNING: restart: lapse = 0 in the T equation, as in SAM (set i_lapse 1 to use lapse0)'
     p%lapse = 0d0
  end if
  p%cfl0 = rv('cfl0'); p%dt0 = rv('dt0'); p%c_smag = rv('c_smag'); p%pr_t = rv('pr_t')

  if (i_les == 4) call read_network()
  rc = sam_create(p, h)
  if (rc /= SAM_OK) call fail(int(rc), 'sam_create: '//trim(sam_error_message()))
  mlp = c_null_ptr
  if (i_les == 4) call attach_network()
  if (i_les == 1) then
     rc = sam_set_cl(h, int(1 + iv('cl_couple'), c_int))
     if (rc /= SAM_OK) call fail(int(rc), 'sam_set_cl: '//trim(sam_error_message()))
  end if
  if (iv('i_force') == 1) then
     rc = sam_set_forcing(h, int(iv('k_force'), c_int))
     if (rc /= SAM_OK) call fail(int(rc), 'sam_set_forcing: '//trim(sam_error_message()))
  end if
  ! planes frames at every nt with mod(nt, n_skip_p) = 0 (n_skip_p defaults
  ! to 1e9: a fresh start still writes time.out and mean.out at nt = 0, as
  ! SAM does); the pressure is kept only when n_skip_p is given
  n_skip_p = iv('n_skip_p')
  planes_on = .true.
  if (found(idx('n_skip_p'))) then
     rc = sam_set_planes(h, 1_c_int)
     if (rc /= SAM_OK) call fail(int(rc), 'sam_set_planes: '//trim(sam_error_message()))
  end if
  rc = sam_get_dims(h, d)
  lu = d%lu
  lm = d%lm
  allocate(u(d%nz_min, d%ny_min, d%nx_min, lu), spec(5 * lm + 3))

  ! ---- initial state: restart files (sam.f:100-108) --------------------
  if (i_restart /= 1) then
     call fail(6, 'i_restart=0: initial_field is not built; write vel.000000/header.000000 '// &
          'with sam_amd.ic and restart from them')
  end if
  write (ext, '(i6.6)') nt_restart
  call read_header('header.'//ext, ierr)
  if (ierr /= 0) call fail(ierr, 'header.'//ext)
  call read_field('vel.'//ext, ierr)
  if (ierr /= 0) call fail(ierr, 'vel.'//ext)
  time = hvalues(idx('time'))
  t_stat = hvalues(idx('t_stat'))
  n_frame_p = nint(hvalues(idx('n_frame_p')))
  nt_start = nint(hvalues(idx('nt')))
  dt = rv('dt0')            ! assign_run_time.f:7-14 keeps the input dt
  rc = sam_set_state(h, u, int(d%nz_min, c_int))
  if (rc /= SAM_OK) call fail(int(rc), 'sam_set_state: '//trim(sam_error_message()))
  rc = sam_set_clock(h, time, dt, int(nt_start, c_long_long))

  appendf = (i_restart == 1)
  call open_files(appendf)
  if (planes_on) call open_planes()

  ! ---- time loop (sam.f:122-326) ---------------------------------------
  nt_end = nt_start + n_steps
  do nt = nt_start, nt_end
     if (planes_on .and. mod(nt, n_skip_p) == 0) call write_planes_out()   ! sam.f:236-239
     want = 0
     if (mod(nt, n_skip_h) == 0 .or. nt == nt_end) want = 1
     if ((nt > 0 .and. mod(nt, n_skip_v) == 0) .or. nt == nt_end) then
        write (ext, '(i6.6)') nt
        if (nt < nt_end .and. iv('vel_midstep') == 1) then
           rc = sam_get_state_substep(h, int(p%nrk_max, c_int), u, int(d%nz_min, c_int))
        else
           rc = sam_get_state(h, u, int(d%nz_min, c_int))
        end if
        if (rc /= SAM_OK) call fail(int(rc), 'vel.'//ext//': '//trim(sam_error_message()))
        call write_header('header.'//ext, time, dt, t_stat, nt)
        call write_field('vel.'//ext)
     end if
     if (nt == nt_end) then
        rc = sam_rhs(h, c_null_ptr, int(d%nz_min, c_int), int(want, c_int), dg)
     else
        rc = sam_step(h, 1_c_int, int(want, c_int), dg)
     end if
     if (rc /= SAM_OK) call fail(int(rc), 'step: '//trim(sam_error_message()))
     dt = dg%dt
     if (want == 1) then
        rc = sam_get_spectra(h, spec)
        call write_logs(nt, time, dg, spec)
     end if
     if (nt /= nt_end) time = dg%time
  end do

  rc = sam_destroy(h)
  if (c_associated(mlp)) rc = sam_mlp_destroy(mlp)
  close (2); close (3); close (4); close (7); close (12)
  if (planes_on) call close_planes()

contains

  ! learned closure for i_les 4: parameters from mlp.params
  subroutine read_network()
    integer :: un, ios
    integer(8) :: fsize
    allocate(mlp_params(SAM_MLP_NPARAM + 32))
    inquire (file='mlp.params', size=fsize)
    if (fsize < 0) call fail(6, 'i_les=4 needs mlp.params (the trained network)')
    if (fsize /= 4_8 * size(mlp_params, kind=8)) &
         call fail(10, 'mlp.params must hold SAM_MLP_NPARAM + 32 FP32 values')
    open (newunit=un, file='mlp.params', access='stream', form='unformatted', status='old', &
         action='read', iostat=ios)
    if (ios == 0) read (un, iostat=ios) mlp_params
    if (ios /= 0) call fail(6, 'cannot read mlp.params')
    close (un)
  end subroutine read_network

  subroutine attach_network()
    rc = sam_mlp_create(mlp, 0_c_int, 0_c_long_long)
    if (rc /= SAM_OK) call fail(int(rc), 'sam_mlp_create: '//trim(sam_error_message()))
    rc = sam_mlp_set_labels(mlp, int(iv('mlp_labels'), c_int), rv('pr_t'))
    if (rc == SAM_OK) rc = sam_mlp_set_activation(mlp, int(iv('mlp_act'), c_int))
    if (rc == SAM_OK) rc = sam_mlp_set_params(mlp, mlp_params)
    if (rc == SAM_OK) rc = sam_attach_mlp(h, mlp)
    if (rc == SAM_OK) rc = sam_set_mlp_refresh(h, int(iv('mlp_refresh'), c_int))
    if (rc /= SAM_OK) call fail(int(rc), 'learned closure: '//trim(sam_error_message()))
  end subroutine attach_network

  subroutine fail(code, what)
    integer, intent(in) :: code
    character(len=*), intent(in) :: what
    print *, 'ERROR (', code, '): ', trim(what)
    error stop code
  end subroutine fail

  integer function idx(name)
    character(len=*), intent(in) :: name
    integer :: k
    idx = 0
    do k = 1, NPAR + NEXT
       if (lower(labels(k)) == lower(name)) then
          idx = k
          return
       end if
    end do
  end function idx

  real(8) function rv(name)
    character(len=*), intent(in) :: name
    rv = values(idx(name))
  end function rv

  integer function iv(name)
    character(len=*), intent(in) :: name
    iv = nint(values(idx(name)))
  end function iv

  function lower(s) result(t)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: k, c
    t = s
    do k = 1, len(s)
       c = ichar(s(k:k))
       if (c >= 65 .and. c <= 90) t(k:k) = achar(c + 32)
    end do
  end function lower

  ! parameter table of set_labels.f:29-102 (+ the three extension labels)
  subroutine set_labels()
    integer :: k, m
    character(len=12) :: r(24), iv_(48)
    real(8) :: rd(24), idf(48)
    logical :: rreq(24), rfix(24), ireq(48), ifix(48)
    r = [character(len=12) :: 'xL', 'yL', 'zL', 'c_smag', 'dt0', 'cfl0', 'k_truncate', 'z0', 'to', &
         'lapse0', 'grav', 'vis', 'pr', 'amplitude', 'lambda_x', 'Gam', 'omega', 'Uo', 'shear', &
         'shear_ratio', 'flct_u', 'flct_t', 'lambda_z_fs', 'Ri_fs']
    rd = [0d0, 0d0, 0d0, 0.01d0, 0.01d0, 1d0, 1d6, 0d0, 1d0, 0d0, 0d0, 0d0, 1d0, 0d0, 0d0, 0d0, &
         0d0, 0d0, 1d0, 2.5d0, 0d0, 0d0, 0d0, 0d0]
    rreq = .false.; rreq(1:3) = .true.; rreq(12) = .true.
    rfix = .true.
    rfix([4, 5, 6, 7, 11, 12, 13, 21, 22, 23, 24]) = .false.
    iv_(1:12) = [character(len=12) :: 'Nx', 'Ny', 'Nz', 'i_restart', 'nt_restart', 'i_strat', 'i_les', &
         'n_steps', 'n_skip_h', 'n_skip_p', 'n_skip_v', 'n_skip_s']
    do m = 1, 9
       write (iv_(12 + m), '(a,i1)') 'k_xy_plane', m
       write (iv_(21 + m), '(a,i1)') 'j_xz_plane', m
       write (iv_(30 + m), '(a,i1)') 'i_yz_plane', m
    end do
    iv_(40:48) = [character(len=12) :: 'i_stat', 'i_cfl', 'nrk_max', 'n_dealias', 'iubc_z', 'i_prob', &
         'i_force', 'k_force', 'i_fs']
    idf = 0d0
    idf(9:12) = 1d9
    idf(41) = 1d0; idf(42) = 3d0; idf(43) = 1d0; idf(45) = 2d0; idf(47) = 2d0; idf(48) = 2d0
    ireq = .false.; ireq(1:3) = .true.; ireq(8) = .true.
    ifix = .false.; ifix(1:3) = .true.; ifix(6) = .true.; ifix(44) = .true.; ifix(45) = .true.
    do k = 1, 24
       labels(k) = r(k); defaults(k) = rd(k); required(k) = rreq(k); fixed(k) = rfix(k)
    end do
    do k = 1, 48
       labels(24 + k) = iv_(k); defaults(24 + k) = idf(k); required(24 + k) = ireq(k); fixed(24 + k) = ifix(k)
    end do
    labels(73:78) = [character(len=12) :: 'time', 'dt', 't_stat', 'nt', 'n_frame_p', 'n_hist']
    defaults(73:78) = 0d0; required(73:78) = .true.; fixed(73:78) = .false.
    labels(79:86) = [character(len=12) :: 'i_prec', 'i_lapse', 'pr_t', 'mlp_labels', 'mlp_refresh', 'cl_couple', &
                     'mlp_act', 'vel_midstep']
    defaults(79:86) = [0d0, 0d0, 1d0, 1d0, 20d0, 0d0, 0d0, 0d0]; required(79:86) = .false.; fixed(79:86) = .false.
  end subroutine set_labels

  ! read_line.f:24-89: one "label value [comment]" per line; '#' in column 1
  ! is a comment; a blank line or the label 'eof' ends the file; c/k/m
  ! suffixes multiply the value by 1e2/1e3/1e6; labels are case-insensitive.
  subroutine read_params(fname, vals, fnd, ierr)
    character(len=*), intent(in) :: fname
    real(8), intent(inout) :: vals(:)
    logical, intent(out) :: fnd(:)
    integer, intent(out) :: ierr
    character(len=256) :: line, label
    integer :: un, ios, b, e, k, c, nb, ne
    real(8) :: val, mult
    logical :: ok
    ierr = 0
    fnd = .false.
    open (newunit=un, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       print *, 'ERROR: cannot open ', trim(fname)
       ierr = 3
       return
    end if
    do
       read (un, '(a)', iostat=ios) line
       if (ios /= 0) exit
       if (line(1:1) == '#') cycle
       if (len_trim(line) == 0) exit
       b = verify(line, ' ')
       e = index(line(b:), ' ') + b - 2
       label = lower(line(b:e))
       ok = .true.
       do k = b, e
          c = ichar(line(k:k))
          if (c < 48 .or. (c > 57 .and. c < 65) .or. (c > 90 .and. c < 95) .or. c == 96 .or. c > 122) ok = .false.
       end do
       if (.not. ok) then
          print *, 'ERROR: Invalid character in label = ', trim(label)
          ierr = 4
          exit
       end if
       if (trim(label) == 'eof') exit
       nb = scan(line(e + 1:), '+-.0123456789')
       if (nb == 0) then
          print *, 'ERROR: No numeric entry found'
          ierr = 5
          exit
       end if
       nb = nb + e
       ne = index(line(nb + 1:), ' ') + nb - 1
       if (ne < nb) ne = len_trim(line)
       mult = 1d0
       select case (line(ne:ne))
       case ('c'); mult = 1d2
       case ('k'); mult = 1d3
       case ('m'); mult = 1d6
       end select
       if (mult /= 1d0) ne = ne - 1
       read (line(nb:ne), *, iostat=ios) val
       if (ios /= 0) then
          print *, 'ERROR: Malformed numeric entry'
          ierr = 2
          exit
       end if
       val = val * mult
       k = idx(trim(label))
       if (k == 0) then
          print *, 'Warning: extraneous variable ', trim(label), ' found in file ', trim(fname)
       else
          fnd(k) = .true.
          vals(k) = val
       end if
    end do
    close (un)
  end subroutine read_params

  ! check_inputs.f:14-231 (planes and GW checks omitted: not part of the step)
  subroutine check_inputs(ierr)
    integer, intent(out) :: ierr
    logical :: req(NPAR + NEXT)
    integer :: k
    ierr = 0
    if (iv('nx') > 2050 .or. iv('ny') > 2050 .or. iv('nz') > 2050) then
       print *, 'ERROR: static dimensions are too small'
       ierr = 7
       return
    end if
    req = required
    req(73:78) = .false.
    if (iv('i_restart') == 1) req(idx('nt_restart')) = .true.
    if (iv('i_les') == 1) req(idx('c_smag')) = .true.
    if (iv('i_cfl') == 0) req(idx('dt0')) = .true.
    if (iv('i_force') == 1) req(idx('k_force')) = .true.
    if (iv('i_strat') == 1) then
       req(idx('lapse0')) = .true.; req(idx('grav')) = .true.; req(idx('z0')) = .true.
       req(idx('to')) = .true.; req(idx('pr')) = .true.
    end if
    if (iv('i_prob') == 5) req(idx('shear')) = .true.
    if (iv('i_fs') /= 0) then
       req(idx('lambda_z_fs')) = .true.; req(idx('ri_fs')) = .true.
    end if
    do k = 1, NIN
       if (req(k) .and. .not. found(k)) then
          print *, 'ERROR: Required input ', trim(labels(k)), ' was not provided'
          ierr = 6
          return
       end if
    end do
    call range_check('i_restart', 0, 1)
    call range_check('i_les', 0, 4)
    call range_check('i_stat', 0, 2)
    call range_check('i_cfl', 0, 2)
    call range_check('nrk_max', 0, 3)
    call range_check('n_dealias', 0, 3)
    call range_check('i_prob', -4, 5)
    call range_check('i_force', 0, 1)
    call range_check('k_force', 0, max(iv('nx') / 2, iv('ny') / 2, iv('nz') / 2))
    call range_check('i_prec', 0, 1)
    call range_check('mlp_labels', 0, 1)
    call range_check('mlp_refresh', 1, 1000000)
    call range_check('cl_couple', 0, 1)
    call range_check('mlp_act', 0, 1)
  contains
    subroutine range_check(name, lo, hi)
      character(len=*), intent(in) :: name
      integer, intent(in) :: lo, hi
      if (ierr /= 0) return
      if (iv(name) < lo .or. iv(name) > hi) then
         print *, 'ERROR: Invalid value for parameter ', trim(name), ' =', rv(name)
         ierr = 9
      end if
    end subroutine range_check
  end subroutine check_inputs

  ! read_header.f:10-42: fixed parameters must agree to 1e-12
  subroutine read_header(fname, ierr)
    character(len=*), intent(in) :: fname
    integer, intent(out) :: ierr
    integer :: k
    hvalues = values
    call read_params(fname, hvalues, hfound, ierr)
    if (ierr /= 0) return
    do k = 1, NIN
       if (.not. hfound(k)) cycle
       if (abs(hvalues(k) - values(k)) > 1d-12) then
          if (fixed(k)) then
             print *, 'ERROR: parameter ', trim(labels(k)), ' differs between the header and input file'
             print *, hvalues(k), values(k), hvalues(k) - values(k)
             ierr = 1
          else if (trim(lower(labels(k))) /= 'i_restart' .and. trim(lower(labels(k))) /= 'nt_restart') then
             print *, 'WARNING: parameter ', trim(labels(k)), ' differs between the header and input file'
          end if
       end if
    end do
    do k = 73, 78
       if (.not. hfound(k)) then
          print *, 'ERROR: header lacks ', trim(labels(k))
          ierr = 6
       end if
    end do
  end subroutine read_header

  ! write_header.f:11-33: list-directed "label value" for every parameter
  subroutine write_header(fname, t, dtv, ts, ntv)
    character(len=*), intent(in) :: fname
    real(8), intent(in) :: t, dtv, ts
    integer, intent(in) :: ntv
    integer :: un, k
    real(8) :: v(NPAR)
    v(1:NIN) = values(1:NIN)
    v(73) = t; v(74) = dtv; v(75) = ts; v(76) = ntv; v(77) = n_frame_p; v(78) = 0
    open (newunit=un, file=fname, status='replace', action='write')
    do k = 1, NPAR
       write (un, *) labels(k), v(k)
    end do
    close (un)
  end subroutine write_header

  ! read_field.f:13-52: raw complex(8) records (Nz_min,Ny_min) per (field, kx)
  subroutine read_field(fname, ierr)
    character(len=*), intent(in) :: fname
    integer, intent(out) :: ierr
    integer :: un, ios
    integer(8) :: expect, fsize
    ierr = 0
    expect = 16_8 * int(d%nz_min, 8) * int(d%ny_min, 8) * int(d%nx_min, 8) * int(lu, 8)
    inquire (file=fname, size=fsize)
    if (fsize /= expect) then
       print *, 'ERROR: Vel file is not of correct size'
       print *, 'isize_expect, isize_found = ', expect, fsize
       ierr = 10
       return
    end if
    open (newunit=un, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
       ierr = 3
       return
    end if
    read (un, iostat=ios) u
    close (un)
    if (ios /= 0) ierr = 10
  end subroutine read_field

  ! write_field.f:12-34
  subroutine write_field(fname)
    character(len=*), intent(in) :: fname
    integer :: un
    open (newunit=un, file=fname, access='stream', form='unformatted', status='replace', action='write')
    write (un) u
    close (un)
  end subroutine write_field

  ! check_inputs.f:27-62 (default first planes at N/2, counts) and
  ! open_files.f:24-58: xy[m].dat direct access (records of Nx*Ny doubles),
  ! xz/yz[m].dat written at byte offsets (mpi_file_write_at there, stream
  ! access here), mean.out direct access (records of Nz*Lu doubles),
  ! time.out sequential from its start
  subroutine open_planes()
    integer :: m
    character(len=1) :: e1
    k_xy = [(int(iv('k_xy_plane'//achar(48 + m)), c_int), m = 1, 9)]
    j_xz = [(int(iv('j_xz_plane'//achar(48 + m)), c_int), m = 1, 9)]
    i_yz = [(int(iv('i_yz_plane'//achar(48 + m)), c_int), m = 1, 9)]
    if (found(idx('n_skip_p'))) then
       if (k_xy(1) == 0) k_xy(1) = p%nz / 2
       if (j_xz(1) == 0) j_xz(1) = p%ny / 2
       if (i_yz(1) == 0) i_yz(1) = p%nx / 2
    end if
    n_xy = count_planes(k_xy)
    n_xz = count_planes(j_xz)
    n_yz = count_planes(i_yz)
    allocate(pmean(p%nz, lu), pxy(p%nx, p%ny, 2 * lu, max(n_xy, 1)), pxz(p%nx, p%nz, 2 * lu, max(n_xz, 1)), &
         pyz(p%ny, p%nz, 2 * lu, max(n_yz, 1)))
    open (unit=10, file='time.out', position='rewind')
    open (unit=11, file='mean.out', form='unformatted', access='direct', recl=p%nz * lu * 8)
    do m = 1, n_xy
       write (e1, '(i1.1)') m
       open (unit=20 + m, file='xy'//e1//'.dat', form='unformatted', access='direct', recl=p%nx * p%ny * 8)
    end do
    do m = 1, n_xz
       write (e1, '(i1.1)') m
       open (unit=30 + m, file='xz'//e1//'.dat', form='unformatted', access='stream')
    end do
    do m = 1, n_yz
       write (e1, '(i1.1)') m
       open (unit=40 + m, file='yz'//e1//'.dat', form='unformatted', access='stream')
    end do
  end subroutine open_planes

  integer function count_planes(a)
    integer(c_int), intent(in) :: a(9)
    integer :: m
    count_planes = 9
    do m = 1, 9
       if (a(m) == 0) then
          count_planes = m - 1
          return
       end if
    end do
  end function count_planes

  ! write_planes.f:18-144 for one frame of the current state
  subroutine write_planes_out()
    integer :: m, n, nf
    integer(8) :: blk, off
    nf = 2 * lu
    rc = sam_planes(h, int(n_xy, c_int), k_xy, int(n_xz, c_int), j_xz, int(n_yz, c_int), i_yz, pmean, pxy, pxz, pyz)
    if (rc /= SAM_OK) call fail(int(rc), 'sam_planes: '//trim(sam_error_message()))
    n_frame_p = n_frame_p + 1
    write (10, '(i6,1p,e12.4)') nt, time
    flush (10)
    write (11, rec=n_frame_p) pmean
    do m = 1, n_xy
       do n = 1, nf
          write (20 + m, rec=(n_frame_p - 1) * nf + n) pxy(:, :, n, m)
       end do
    end do
    blk = int(p%nx, 8) * int(p%nz, 8) * int(nf * 8, 8)
    do m = 1, n_xz
       do n = 1, nf
          off = int(n_frame_p - 1, 8) * blk + int(n - 1, 8) * int(p%nx, 8) * int(p%nz, 8) * 8
          write (30 + m, pos=off + 1) pxz(:, :, n, m)
       end do
    end do
    blk = int(p%ny, 8) * int(p%nz, 8) * int(nf * 8, 8)
    do m = 1, n_yz
       do n = 1, nf
          off = int(n_frame_p - 1, 8) * blk + int(n - 1, 8) * int(p%ny, 8) * int(p%nz, 8) * 8
          write (40 + m, pos=off + 1) pyz(:, :, n, m)
       end do
    end do
  end subroutine write_planes_out

  subroutine close_planes()
    integer :: m
    close (10); close (11)
    do m = 1, n_xy
       close (20 + m)
    end do
    do m = 1, n_xz
       close (30 + m)
    end do
    do m = 1, n_yz
       close (40 + m)
    end do
  end subroutine close_planes

  ! open_files.f:90-130 (ascii logs, appended on restart)
  subroutine open_files(app)
    logical, intent(in) :: app
    character(len=8) :: pos
    pos = 'rewind'
    if (app) pos = 'append'
    open (unit=2, file='energy.out', position=pos)
    open (unit=3, file='umax.out', position=pos)
    open (unit=4, file='spectra.out', position=pos)
    open (unit=7, file='tstep.out', position=pos)
    open (unit=12, file='variance.out', position=pos)
    if (.not. app) then
       write (2, '(a)') '#   nt    time       energy'
       write (3, '(a)') '#   nt    time       div_max      u_max       v_max       w_max       T_max'
       write (7, '(a)') '#   nt    time         dt         cfl_x       cfl_y       cfl_z      cfl_vis'
       write (12, '(a)') '#   nt    time        u_var       v_var       w_var      0.5u_sq      T_var'
    end if
  end subroutine open_files

  ! sam.f:248-273 and spectra.f:92-111
  subroutine write_logs(ntv, t, g, sp)
    integer, intent(in) :: ntv
    real(8), intent(in) :: t
    type(sam_diag), intent(in) :: g
    real(c_double), intent(in) :: sp(:)
    integer :: k
    real(8) :: pi, sumdk
    pi = acos(-1d0)
    print '(i6,1p11e12.4)', ntv, t, 0.5d0 * (g%u_var(1) + g%u_var(2) + g%u_var(3))
    write (2, '(i6,1p11e12.4)') ntv, t, sp(5 * lm + 1)
    write (3, '(i6,1p11e12.4)') ntv, t, sp(5 * lm + 3), g%u_max(1:lu)
    write (7, '(i6,1p11e12.4)') ntv, t, g%dt, g%cfl_x, g%cfl_y, g%cfl_z, 0d0
    write (12, '(i6,1p11e12.4)') ntv, t, g%u_var(1:3), 0.5d0 * (g%u_var(1) + g%u_var(2) + g%u_var(3)), &
         g%u_var(4:lu)
    sumdk = sum(sp(2 * lm + 1:3 * lm))
    write (4, 10) ntv, t, sp(5 * lm + 1), sumdk
10  format('nt =', i3, '  t =', 1pe11.4, '  KE =', 1pe11.4, '  Diss =', 1pe11.4, //, '           k     E(k)')
    do k = 1, lm
       if (sp(4 * lm + k) /= 0d0) then
          write (4, '(i12,1p,4e12.4)') k - 1, 4d0 * pi * ((k - 1)**2) / sp(4 * lm + k) * (0.5d0 * sp(k)), &
               sp(lm + k) / sp(4 * lm + k), sp(2 * lm + k) / sp(4 * lm + k), sp(3 * lm + k) / sp(4 * lm + k)
       end if
    end do
    write (4, *)
    flush (2); flush (3); flush (4); flush (7); flush (12)
  end subroutine write_logs

end program samhip_driver
