!------------------------------------------------------------------------------!
! Driver of the MI355X build: the reference's main program (main.f90) with the   !
! per-step work delegated to the blgpu C ABI through ISO_C_BINDING.             !
!                                                                              !
! Kept from the reference, on the host: the .turbb reader (params.f90,          !
! miscel.f90), init_flow and the snapshot reader (input_output.f90:130-487),    !
! Blasius / temporal-mode boundary tables (boundary_conditions.f90:1452-1797),  !
! the inlet thickness and the statistics text file (statistics.f90:177-201,     !
! input_output.f90:685-746), the monitor block (monitor.f90:20-173) and the     !
! snapshot writer (input_output.f90:496-680).  The z means of the statistics   !
! come from the device (bl_statistics); snapshot fields are copied to the host  !
! behind the next step (bl_snapshot_begin / bl_snapshot_end).  Single rank     !
! only, as the reference's GPU path (initialization.f90:311).                  !
!------------------------------------------------------------------------------!
module bl_params
  use, intrinsic :: iso_fortran_env, only : int32, real64
  implicit none
  character(len=512) :: paramsfilename = ' '
contains
  ! getvar (miscel.f90:32-70): first line whose text before '=' is the key;
  ! the value is what follows '=', cut at '!'
  subroutine getvar(name, str)
    character(len=*), intent(in)  :: name
    character(len=*), intent(out) :: str
    character(len=1024) :: line
    integer :: iu, ios, eq, bang
    logical :: ex
    str = ' '
    inquire(file=trim(paramsfilename), exist=ex)
    if (.not. ex) return
    open(newunit=iu, file=trim(paramsfilename), status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do
       read(iu, '(A)', iostat=ios) line
       if (ios /= 0) exit
       eq = scan(line, '=')
       if (eq < 2) cycle
       if (trim(adjustl(line(1:eq-1))) == name) then
          str = line(eq+1:)
          bang = scan(str, '!')
          if (bang /= 0) str = str(1:bang-1)
          exit
       end if
    end do
    close(iu)
  end subroutine getvar

  subroutine get_dbl(name, v, found)
    character(len=*), intent(in) :: name
    real(real64), intent(inout) :: v
    logical, intent(out) :: found
    character(len=1024) :: s
    call getvar(name, s)
    found = len_trim(s) > 0
    if (found) read(s, *) v
  end subroutine get_dbl

  subroutine get_int(name, v, found)
    character(len=*), intent(in) :: name
    integer, intent(inout) :: v
    logical, intent(out) :: found
    character(len=1024) :: s
    call getvar(name, s)
    found = len_trim(s) > 0
    if (found) read(s, *) v
  end subroutine get_int

  subroutine get_str(name, v, found)
    character(len=*), intent(in) :: name
    character(len=*), intent(inout) :: v
    logical, intent(out) :: found
    character(len=1024) :: s
    call getvar(name, s)
    found = len_trim(s) > 0
    if (found) v = trim(adjustl(s))
  end subroutine get_str

  subroutine get_int_arr(name, v, found)
    character(len=*), intent(in) :: name
    integer, intent(inout) :: v(:)
    logical, intent(out) :: found
    character(len=1024) :: s
    call getvar(name, s)
    found = len_trim(s) > 0
    if (found) read(s, *) v
  end subroutine get_int_arr

  subroutine get_dbl_arr(name, v, found)
    character(len=*), intent(in) :: name
    real(real64), intent(inout) :: v(:)
    logical, intent(out) :: found
    character(len=1024) :: s
    call getvar(name, s)
    found = len_trim(s) > 0
    if (found) read(s, *) v
  end subroutine get_dbl_arr
end module bl_params


module bl_state
  use, intrinsic :: iso_fortran_env, only : int32, int64, real64
  implicit none
  ! parameters (params.f90:119-198)
  integer :: nx, ny, nz, nxg, nyg, nzg, nxm, nym, nzm
  integer :: nsteps = 0, nsave = 1, nstats = 1, nmonitor = 1, itime_step = 3, les_model = 0
  integer :: nut_bc = 0   ! Dirichlet_nu_t (initialization.f90:586, params.f90:156)
  ! optional key of this build (absent from the reference's files): 0 = FP64,
  ! 1 = mixed precision (FP32 fields, split-BF16x3 transforms, FP64 Thomas)
  integer :: iprecision = 0
  integer :: iwall_model = 0, inflow_flag = 1, top_flag = 0, random_init = 1
  ! boundary-condition parameters (params.f90:122-125, 144, 162); beta_hartree
  ! (Falkner-Skan) is 0 in the reference (input_output.f90:36) and an optional
  ! key of this build
  real(real64) :: vbs_max = 0d0, x_bs = 0d0, sigma_bs = 1d0, phi_bs = 0d0, amplitude = 0d0, beta_hartree = 0d0
  integer :: i_rescale = 0
  ! Lund rescaling (params.f90:144-146)
  real(real64) :: delta_inlet = 0d0, t_resc = 0d0
  ! wall model (params.f90:154-155, 198-206): TauwModel, and alphas =
  ! alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult
  integer :: istress_model = 0
  real(real64) :: alphas(5) = 0d0
  ! models 9, 12: WMnutflag (iwall_model_nut) and WMnut (frac_vis_wall_model,
  ! params.f90:159-160)
  integer :: iwall_model_nut = 0
  real(real64) :: frac_vis = 0d0
  integer :: nstep_init = 0, nstep_init_input = -45, istep = 0
  real(real64) :: nu = 0d0, cfl = -1d-4, dPdx = 0d0, dPdz = 0d0, t = 0d0, dt = 0d0
  real(real64) :: boxsize(4) = 0d0
  character(len=512) :: file_inflow = ' ', file_temporal = ' ', filein = ' ', fileout = ' '
  ! grids (initialization.f90:174-227)
  real(real64), allocatable, target :: x(:), y(:), z(:), xm(:), ym(:), zm(:), xg(:), yg(:), zg(:)
  real(real64), allocatable :: wy0(:), wy1(:)
  ! fields, reference shapes
  real(real64), allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), P(:,:,:), nu_t(:,:,:)
  ! boundary tables
  real(real64), allocatable, target :: U_inlet(:), V_inlet(:), W_inlet(:), U_top(:), V_top(:), W_top(:)
  integer :: n_modes = 0, m_modes = 0
  real(real64), allocatable, target :: zmode(:), tmode(:)
  real(real64), allocatable, target :: qu_re(:,:,:), qu_im(:,:,:), qv_re(:,:,:), qv_im(:,:,:)
  real(real64), allocatable, target :: qw_re(:,:,:), qw_im(:,:,:)
  ! statistics (statistics.f90)
  real(real64), allocatable :: Cf(:), Umean(:,:), Vmean(:,:), Wmean(:,:), U2mean(:,:), V2mean(:,:)
  real(real64), allocatable :: W2mean(:,:), UVmean(:,:), Pmean(:,:), P2mean(:,:)
  real(real64) :: Retheta_inlet = 0d0, Redelta_inlet = 0d0, Rex_inlet = 0d0, delta99_inlet = 0d0
  ! Blasius table (text file: n, eta(n), f(n), f'(n))
  integer :: n_source = 0
  real(real64), allocatable :: eta_s(:), f_s(:), df_s(:)
end module bl_state


module bl_setup
  use bl_state
  use bl_params
  implicit none
contains

  subroutine read_parameters()
    integer :: i, ist, ics, nxyz(3)
    logical :: f, f1
    character(len=512) :: arg
    ! -i <file> (params.f90:96-110)
    f = .false.
    do i = 1, command_argument_count()
       call get_command_argument(i, arg)
       if (trim(arg) == '-i') then
          call get_command_argument(i + 1, paramsfilename)
          f = .true.
       end if
    end do
    if (.not. f) stop 'No input file provided. Use *.exe -i <input_file>'
    inquire(file=trim(paramsfilename), exist=f)
    if (.not. f) then
       write(*,*) 'input file: ', trim(paramsfilename), ' does not exist'
       stop
    end if
    call get_dbl('nu', nu, f)
    call get_dbl('CFL', cfl, f)
    call get_int('nsteps', nsteps, f)
    call get_int('nsave', nsave, f)
    call get_int('nstats', nstats, f)
    call get_int('nmonitor', nmonitor, f)
    nxyz = 0
    call get_int_arr('nxyz', nxyz, f)
    call get_dbl_arr('boxsize', boxsize, f)
    call get_int('inflow_flag', inflow_flag, f)
    call get_int('top_flag', top_flag, f)
    call get_dbl('Vmax', vbs_max, f)
    call get_dbl('x0', x_bs, f)
    call get_dbl('sigma', sigma_bs, f)
    call get_dbl('phi', phi_bs, f)
    call get_int('Lund_ix', i_rescale, f)
    call get_dbl('Lund_deltai', delta_inlet, f)
    call get_dbl('Lund_T', t_resc, f)
    call get_dbl('Amplitude', amplitude, f)
    call get_dbl('beta_hartree', beta_hartree, f)
    call get_str('inflow_file', file_inflow, f)
    if (.not. f) stop ' ERROR: you must specify a BL profile in inflow_file '
    call get_dbl('dPdx', dPdx, f)
    call get_dbl('dPdz', dPdz, f)
    call get_int('LES', les_model, f)
    call get_int('nutBC', nut_bc, f)
    call get_int('precision', iprecision, f)
    call get_int('WM', iwall_model, f)
    call get_int('TauwModel', istress_model, f)
    call get_dbl_arr('alphas', alphas, f)
    call get_int('WMnutflag', iwall_model_nut, f)
    call get_dbl('WMnut', frac_vis, f)
    call get_int('init_step', nstep_init_input, f)
    call get_int('init_rand', random_init, f)
    call get_str('filein', filein, f1)
    if (random_init == 1 .and. f1) stop ' ERROR: init_rand = 1, but filein exists'
    if (nstep_init_input /= -45) nstep_init = nstep_init_input
    call get_str('fileout', fileout, f)
    ! create the folder of fileout if needed (params.f90:158-176)
    i = index(fileout, '/', back=.true.)
    if (i > 1) call execute_command_line('mkdir -p ' // trim(adjustl(fileout(1:i-1))), exitstat=ist, cmdstat=ics)
    call get_str('timeinflow_file', file_temporal, f)
    call get_int('RKscheme', itime_step, f)
    nx = nxyz(1); ny = nxyz(2); nz = nxyz(3)
    ! this build runs the RK path of the reference with the GPU boundary
    ! conditions (time_integration.f90:145-465)
    if (itime_step /= 2 .and. itime_step /= 3) stop 'Error: only RKscheme = 2 or 3 is built'
    if (all(iwall_model /= [0, 4, 5, 6, 8, 9, 11, 12, 15])) stop 'Error: WM must be 0, 4, 5, 6, 8, 9, 11, 12 or 15'
    if (les_model < 0 .or. les_model > 3) stop 'Error: LES must be 0, 1, 2 or 3'
    if (iprecision /= 0 .and. iprecision /= 1) stop 'Error: precision must be 0 (FP64) or 1 (mixed)'
    if (inflow_flag < 1 .or. inflow_flag > 5) stop 'Error: inflow_flag must be 1 to 5'
    if ((inflow_flag == 3 .or. inflow_flag == 5) .and. random_init == 1) &
         stop 'Error: Lund inflow (3, 5) starts from a flow field file (init_rand = 0)'
    if (top_flag < 0 .or. top_flag > 3) stop 'Error: top_flag must be 0, 1, 2 or 3'
  end subroutine read_parameters

  subroutine allocate_all()
    nxg = nx + 1; nyg = ny + 1; nzg = nz + 1
    nxm = nx - 1; nym = ny - 1; nzm = nz - 1
    allocate(x(nx), y(ny), z(nz), xm(nxm), ym(nym), zm(nzm), xg(nxg), yg(nyg), zg(nzg))
    allocate(U(nx, nyg, nzg), V(nxg, ny, nzg), W(nxg, nyg, nz), P(nxg, nyg, nzg), nu_t(nxg, nyg, nzg))
    U = 0d0; V = 0d0; W = 0d0; P = 0d0; nu_t = 0d0
    allocate(wy0(ny), wy1(ny))
    allocate(U_inlet(nyg), V_inlet(ny), W_inlet(nyg), U_top(nx), V_top(nxg), W_top(nxg))
    allocate(Cf(nx), Umean(nx, ny), Vmean(nx, ny), Wmean(nx, ny), U2mean(nx, ny), V2mean(nx, ny))
    allocate(W2mean(nx, ny), UVmean(nx, ny), Pmean(nx, ny), P2mean(nx, ny))
  end subroutine allocate_all

  ! compute_turbulent_solution_for_bc (boundary_conditions.f90:1371-1425):
  ! U_top = 1, V_top from the 1/7-power-law growth of delta, W_top = 0; for
  ! inflow_flag 5 the inlet profile U_inlet(nyg), V_inlet(ny) from the
  ! inflow_file stream (an int32 count ny, then the two arrays), W_inlet = 0
  subroutine turbulent_tables()
    integer :: iu, ny_in, i
    real(real64) :: rex
    U_top = 1d0
    do i = 1, nxg
       rex = 1d0 * xg(i) / nu
       V_top(i) = 1d0 * (0.0160d0 * rex**(-1d0 / 7d0))
    end do
    W_top = 0d0
    U_inlet = 0d0; V_inlet = 0d0; W_inlet = 0d0
    n_modes = 0; m_modes = 0
    if (inflow_flag == 5) then
       write(*,*) 'Reading Turbulent profile for inlet from file'
       open(newunit=iu, file=trim(file_inflow), form='unformatted', access='stream', action='read', status='old')
       read(iu) ny_in
       if (ny_in /= ny) stop 'Profile doesnt match dimensions'
       read(iu) U_inlet
       read(iu) V_inlet
       close(iu)
    end if
  end subroutine turbulent_tables

  subroutine read_blasius_table()
    integer :: iu
    open(newunit=iu, file=trim(file_inflow), form='formatted', action='read', status='old')
    read(iu, *) n_source
    allocate(eta_s(n_source), f_s(n_source), df_s(n_source))
    read(iu, *) eta_s
    read(iu, *) f_s
    read(iu, *) df_s
    close(iu)
  end subroutine read_blasius_table

  ! init_flow grid (input_output.f90:150-174); the Blasius field itself is
  ! built by bl_blasius_field with the same interpolation (:206-259)
  subroutine init_flow_grid()
    integer :: i
    real(real64) :: a
    do i = 1, nx
       x(i) = real(i - 1, real64)
    end do
    x = x - x(1)
    x = 1d0 + x / maxval(x) * boxsize(1)
    do i = 1, nz
       z(i) = real(i - 1, real64)
    end do
    z = z / z(nz - 1) * boxsize(3)
    do i = 1, ny
       y(i) = real(i - 1, real64)
    end do
    y = y - y(1)
    y = y / maxval(y)
    a = boxsize(4)
    do i = 1, ny
       y(i) = dsinh(a * y(i)) / dsinh(a)
    end do
    y = y - y(1)
    y = y / maxval(y) * boxsize(2)
    t = 0d0
  end subroutine init_flow_grid

  ! read_input_data, single rank (input_output.f90:288-487): header, then
  ! U, V, W at full allocated extents located by offsets from the header end
  subroutine read_input_data()
    integer :: iu, nn(3), nxf, nyf, nzf, ndum
    integer(int64) :: pos_header, nsize_U, nsize_V
    real(real64) :: nu_dummy
    open(newunit=iu, file=trim(filein), access='stream', form='unformatted', action='read', status='old')
    read(iu) t, nu_dummy
    read(iu) nxf
    if (nxf == -73) then
       read(iu) ndum
       if (nstep_init_input == -45) then
          nstep_init = ndum
          write(*,*) 'Reading init step from file: ', nstep_init
       end if
       read(iu) nxf
    end if
    if (nxf /= nx) stop 'nx in file does not match nxyz'
    read(iu) x
    read(iu) nyf
    if (nyf /= ny) stop 'ny in file does not match nxyz'
    read(iu) y
    read(iu) nzf
    if (nzf /= nz) stop 'nz in file does not match nxyz'
    read(iu) z
    read(iu) ndum
    read(iu) xm
    read(iu) ndum
    read(iu) ym
    read(iu) ndum
    read(iu) zm
    inquire(iu, pos=pos_header)
    pos_header = pos_header - 1
    nsize_U = int(nx, int64) * nyg * nzg * 8_int64
    nsize_V = int(nxg, int64) * ny * nzg * 8_int64
    read(iu) nn
    read(iu) U
    read(iu, pos=pos_header + 3*4 + nsize_U + 1) nn
    read(iu) V
    read(iu, pos=pos_header + 3*4 + nsize_U + 3*4 + nsize_V + 1) nn
    read(iu) W
    close(iu)
    if (any(U /= U)) stop 'Error U NaNs!'
    if (any(V /= V)) stop 'Error V NaNs!'
    if (any(W /= W)) stop 'Error W NaNs!'
  end subroutine read_input_data

  ! initialization.f90:181-222, 482-484
  subroutine build_grids()
    integer :: i
    do i = 1, nxm
       xm(i) = 0.5d0 * (x(i) + x(i + 1))
    end do
    do i = 1, nym
       ym(i) = 0.5d0 * (y(i) + y(i + 1))
    end do
    do i = 1, nzm
       zm(i) = 0.5d0 * (z(i) + z(i + 1))
    end do
    xg(2:nxm+1) = xm
    xg(1) = xm(1) - 2d0 * (xm(1) - x(1))
    xg(nxm+2) = xm(nxm) + 2d0 * (x(nx) - xm(nxm))
    yg(2:nym+1) = ym
    yg(1) = ym(1) - 2d0 * (ym(1) - y(1))
    yg(nym+2) = ym(nym) + 2d0 * (y(ny) - ym(nym))
    zg(2:nzm+1) = zm
    zg(1) = zm(1) - 2d0 * (zm(1) - z(1))
    zg(nzm+2) = zm(nzm) + 2d0 * (z(nz) - zm(nzm))
    wy0 = (yg(2:nyg) - y(1:ny)) / (yg(2:nyg) - yg(1:nyg-1))
    wy1 = 1d0 - wy0
  end subroutine build_grids

  ! apply_periodic_bc_z, single rank (boundary_conditions.f90:739-757)
  subroutine periodic_z()
    U(:,:,1) = U(:,:,nzg-2); U(:,:,nzg-1) = U(:,:,2); U(:,:,nzg) = U(:,:,3)
    V(:,:,1) = V(:,:,nzg-2); V(:,:,nzg-1) = V(:,:,2); V(:,:,nzg) = V(:,:,3)
    W(:,:,1) = W(:,:,nz-1);  W(:,:,nz) = W(:,:,2)
  end subroutine periodic_z

  ! temporal inflow modes: read (boundary_conditions.f90:1667-1732) and
  ! interpolate to yg / y (:1734-1776)
  subroutine read_temporal_modes()
    integer :: iu, ny_in, j, jj, j0, j1
    real(real64) :: beta, omega
    real(real64), allocatable :: ym_in(:), qr(:,:,:), qi(:,:,:)
    if (len_trim(file_temporal) == 0 .or. file_temporal(1:4) == '**TO') then
       write(*,*) 'No temporal inflow file -> pure Blasius (no perturbations)'
       n_modes = 0; m_modes = 0
       return
    end if
    open(newunit=iu, file=trim(file_temporal), form='unformatted', access='stream', action='read', status='old')
    read(iu) ny_in
    read(iu) n_modes
    read(iu) m_modes
    read(iu) beta
    read(iu) omega
    allocate(ym_in(ny_in), zmode(n_modes), tmode(m_modes))
    allocate(qr(ny_in, n_modes, m_modes), qi(ny_in, n_modes, m_modes))
    read(iu) ym_in
    read(iu) zmode
    read(iu) tmode
    zmode = beta * zmode
    tmode = omega * tmode
    allocate(qu_re(nyg, n_modes, m_modes), qu_im(nyg, n_modes, m_modes))
    allocate(qv_re(ny, n_modes, m_modes), qv_im(ny, n_modes, m_modes))
    allocate(qw_re(nyg, n_modes, m_modes), qw_im(nyg, n_modes, m_modes))
    read(iu) qr
    read(iu) qi
    call interp(qr, qi, yg, qu_re, qu_im)
    read(iu) qr
    read(iu) qi
    call interp(qr, qi, y, qv_re, qv_im)
    read(iu) qr
    read(iu) qi
    call interp(qr, qi, yg, qw_re, qw_im)
    close(iu)
  contains
    subroutine interp(sr, si, ygrid, orr, oii)
      real(real64), intent(in) :: sr(:,:,:), si(:,:,:), ygrid(:)
      real(real64), intent(out) :: orr(:,:,:), oii(:,:,:)
      do j = 1, size(ygrid)
         j1 = 0
         do jj = 2, ny_in
            if (ym_in(jj) >= ygrid(j)) then
               j1 = jj
               exit
            end if
            if (jj == ny_in) j1 = ny_in
         end do
         if (j1 == ny_in) then
            orr(j,:,:) = 0d0
            oii(j,:,:) = 0d0
         else
            j0 = j1 - 1
            orr(j,:,:) = sr(j0,:,:) + (ygrid(j) - ym_in(j0)) * (sr(j1,:,:) - sr(j0,:,:)) / (ym_in(j1) - ym_in(j0))
            oii(j,:,:) = si(j0,:,:) + (ygrid(j) - ym_in(j0)) * (si(j1,:,:) - si(j0,:,:)) / (ym_in(j1) - ym_in(j0))
         end if
      end do
    end subroutine interp
  end subroutine read_temporal_modes

end module bl_setup


module bl_output
  use bl_state
  implicit none
contains

  ! compute_statistics (statistics.f90:44-201): the z means and Cf from the
  ! device (bl_statistics), the inlet thickness on the host
  subroutine compute_statistics(h)
    use, intrinsic :: iso_c_binding, only : c_ptr, c_null_ptr
    use blgpu_api
    type(c_ptr), intent(in) :: h
    integer :: j, jref
    real(real64) :: theta, Umean99, w0, w1
    real(real64) :: temp(ny)
    real(real64), allocatable :: means(:,:,:)
    allocate(means(nx, ny, 9))
    call bl_check(bl_statistics(h, means, c_null_ptr, Cf), 'bl_statistics')
    Umean = means(:, :, 1); Vmean = means(:, :, 2); Wmean = means(:, :, 3)
    U2mean = means(:, :, 4); V2mean = means(:, :, 5); W2mean = means(:, :, 6)
    UVmean = means(:, :, 7); Pmean = means(:, :, 8); P2mean = means(:, :, 9)
    deallocate(means)
    ! momentum thickness and delta99 at the inlet (statistics.f90:177-201)
    theta = 0d0
    temp = Umean(1, :) * (1d0 - Umean(1, :))
    do j = 2, ny
       theta = theta + 0.5d0 * (temp(j) + temp(j - 1)) * (y(j) - y(j - 1))
    end do
    Umean99 = 0.99d0
    jref = 0
    do j = 1, ny
       if (Umean(1, j) >= Umean99) then
          jref = j
          exit
       end if
    end do
    if (jref < 2) jref = 2
    w1 = (Umean99 - Umean(1, jref - 1)) / (Umean(1, jref) - Umean(1, jref - 1))
    w0 = 1d0 - w1
    delta99_inlet = w1 * y(jref) + w0 * y(jref - 1)
    Rex_inlet = x(1) / nu
    Retheta_inlet = theta / nu
    Redelta_inlet = delta99_inlet / nu
  end subroutine compute_statistics

  ! output_statistics (input_output.f90:685-746)
  subroutine output_statistics()
    character(len=512) :: fname, fmt
    character(len=8) :: ext, ext_nx, ext_ny
    integer :: iu, ii
    write(ext, '(I0.8)') istep + nstep_init
    write(ext_nx, '(I8)') nx
    write(ext_ny, '(I8)') ny
    fname = trim(adjustl(fileout)) // '.' // trim(adjustl(ext)) // '.stats.txt'
    write(*,*) 'writting ', trim(adjustl(fname))
    open(newunit=iu, file=fname, form='formatted', action='write')
    write(iu, '(A,2F15.8,4I8)') '%', t, nu, nx, ny, nz, istep
    fmt = '(' // trim(adjustl(ext_nx)) // 'F15.8)'
    write(iu, fmt) Cf
    write(iu, fmt) x
    fmt = '(' // trim(adjustl(ext_ny)) // 'F15.8)'
    write(iu, fmt) y
    do ii = 1, nx
       write(iu, fmt) Umean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) Vmean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) Wmean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) U2mean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) V2mean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) W2mean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) Pmean(ii, :)
    end do
    do ii = 1, nx
       write(iu, fmt) P2mean(ii, :)
    end do
    close(iu)
  end subroutine output_statistics

  ! output_data, single rank (input_output.f90:496-666): the z extents are
  ! one plane short for U, V, W and nu_t, P is written in full; the fields
  ! of step istep_s at time t_s
  subroutine output_data(Us, Vs, Ws, Ps, NTs, istep_s, t_s)
    real(real64), intent(in) :: Us(nx, nyg, nzg), Vs(nxg, ny, nzg), Ws(nxg, nyg, nz)
    real(real64), intent(in) :: Ps(nxg, nyg, nzg), NTs(nxg, nyg, nzg)
    integer, intent(in) :: istep_s
    real(real64), intent(in) :: t_s
    character(len=512) :: fname, lnk
    character(len=8) :: ext
    integer :: iu, ist, ics
    logical :: ex
    write(ext, '(I0.8)') istep_s + nstep_init
    fname = trim(adjustl(fileout)) // '.' // trim(adjustl(ext))
    write(*,*) 'writting ', trim(adjustl(fname))
    open(newunit=iu, file=fname, access='stream', form='unformatted', action='write', status='replace')
    write(iu) t_s, nu
    write(iu) -73, istep_s + nstep_init
    write(iu) shape(x), x
    write(iu) shape(y), y
    write(iu) shape(z), z
    write(iu) shape(xm), xm
    write(iu) shape(ym), ym
    write(iu) shape(zm), zm
    write(iu) nx, nyg, nzg
    write(iu) Us(:,:,1:nzg-1)
    write(iu) nxg, ny, nzg
    write(iu) Vs(:,:,1:nzg-1)
    write(iu) nxg, nyg, nz
    write(iu) Ws(:,:,1:nz-1)
    write(iu) nxg, nyg, nzg
    write(iu) NTs(:,:,1:nzg-1)
    write(iu) nxg, nyg, nzg
    write(iu) Ps
    close(iu)
    ! restart symlink (input_output.f90:651-664); like the reference's
    ! `call system`, a failing command is not an error
    lnk = trim(adjustl(fileout)) // '.restart'
    inquire(file=trim(lnk), exist=ex)
    if (ex) call execute_command_line('rm ' // trim(lnk), exitstat=ist, cmdstat=ics)
    call execute_command_line('ln -s ' // trim(adjustl(fname)) // ' ' // trim(lnk), exitstat=ist, cmdstat=ics)
  end subroutine output_data

  ! the fields of the snapshot begun at step istep_s: wait for their copy
  ! (it ran under the kernels queued since) and write them
  subroutine write_snapshot(h, istep_s, t_s)
    use, intrinsic :: iso_c_binding, only : c_ptr, c_f_pointer
    use blgpu_api
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: istep_s
    real(real64), intent(in) :: t_s
    type(c_ptr) :: pu, pv, pw, pp, pn
    real(real64), pointer :: Us(:,:,:), Vs(:,:,:), Ws(:,:,:), Ps(:,:,:), NTs(:,:,:)
    call bl_check(bl_snapshot_end(h, pu, pv, pw, pp, pn), 'bl_snapshot_end')
    call c_f_pointer(pu, Us, [nx, nyg, nzg])
    call c_f_pointer(pv, Vs, [nxg, ny, nzg])
    call c_f_pointer(pw, Ws, [nxg, nyg, nz])
    call c_f_pointer(pp, Ps, [nxg, nyg, nzg])
    call c_f_pointer(pn, NTs, [nxg, nyg, nzg])
    call output_data(Us, Vs, Ws, Ps, NTs, istep_s, t_s)
  end subroutine write_snapshot

end module bl_output


program bl_driver
  use, intrinsic :: iso_c_binding
  use bl_state
  use bl_setup
  use bl_output
  use blgpu_api
  implicit none
  type(bl_config), target :: cfg
  type(bl_boundary) :: bnd
  type(bl_wall_params) :: wmp
  type(c_ptr) :: h
  real(c_double) :: mon(8), div, t_dev
  real(real64) :: time1, time2, snap_t
  integer(int64) :: c0, c1, crate
  integer :: i, snap_step
  logical :: snap_pending

  write(*,*) '----------------------------------------------------------------------'
  write(*,*) '  boundary layer, MI355X per-step path (blgpu) with the reference driver'
  write(*,*) '----------------------------------------------------------------------'
  call read_parameters()
  call allocate_all()
  if (inflow_flag /= 3 .and. inflow_flag /= 5) call read_blasius_table()
  if (random_init == 1) then
     call init_flow_grid()
  else
     call read_input_data()
  end if
  call build_grids()
  if (random_init == 1) then
     call bl_check(bl_blasius_field(nx, ny, nz, x, y, nu, n_source, eta_s, f_s, df_s, U, V, W), &
                   'bl_blasius_field')
  end if
  call periodic_z()
  if (inflow_flag == 3 .or. inflow_flag == 5) then
     call turbulent_tables()
  else
     call bl_check(bl_blasius_tables(nx, ny, x, y, nu, n_source, eta_s, f_s, df_s, U_inlet, V_inlet, &
                   W_inlet, U_top, V_top, W_top), 'bl_blasius_tables')
     call read_temporal_modes()
  end if

  cfg%nx = nx; cfg%ny = ny; cfg%nz = nz
  cfg%x = c_loc(x); cfg%y = c_loc(y); cfg%z = c_loc(z)
  cfg%nu = nu; cfg%dPdx = dPdx; cfg%dPdz = dPdz; cfg%cfl = cfl
  cfg%scheme = itime_step; cfg%les = les_model; cfg%precision = merge(BL_PREC_MIXED, BL_PREC_FP64, iprecision == 1)
  cfg%nut_bc = nut_bc
  cfg%U_inlet = c_loc(U_inlet); cfg%V_inlet = c_loc(V_inlet); cfg%W_inlet = c_loc(W_inlet)
  cfg%U_top = c_loc(U_top); cfg%V_top = c_loc(V_top); cfg%W_top = c_loc(W_top)
  cfg%n_modes = n_modes; cfg%m_modes = m_modes
  cfg%zmode = c_null_ptr; cfg%tmode = c_null_ptr
  cfg%qu_re = c_null_ptr; cfg%qu_im = c_null_ptr; cfg%qv_re = c_null_ptr
  cfg%qv_im = c_null_ptr; cfg%qw_re = c_null_ptr; cfg%qw_im = c_null_ptr
  if (n_modes > 0) then
     cfg%zmode = c_loc(zmode); cfg%tmode = c_loc(tmode)
     cfg%qu_re = c_loc(qu_re); cfg%qu_im = c_loc(qu_im)
     cfg%qv_re = c_loc(qv_re); cfg%qv_im = c_loc(qv_im)
     cfg%qw_re = c_loc(qw_re); cfg%qw_im = c_loc(qw_im)
  end if
  call bl_check(bl_create(cfg, h), 'bl_create')
  call bl_check(bl_upload(h, BL_FIELD_U, U), 'bl_upload U')
  call bl_check(bl_upload(h, BL_FIELD_V, V), 'bl_upload V')
  call bl_check(bl_upload(h, BL_FIELD_W, W), 'bl_upload W')
  call bl_check(bl_upload(h, BL_FIELD_P, P), 'bl_upload P')
  call bl_check(bl_upload(h, BL_FIELD_NUT, nu_t), 'bl_upload nu_t')
  call bl_check(bl_set_time(h, t), 'bl_set_time')
  ! the general dispatcher's inflow and top flags (boundary_conditions.f90:22-121),
  ! after the fields: Lund rescaling starts its time averages from them
  if (inflow_flag /= 1 .or. top_flag /= 0) then
     bnd%inflow_flag = inflow_flag
     bnd%top_flag = top_flag
     bnd%amplitude = amplitude
     bnd%seed = 0_c_long_long
     bnd%vbs_max = vbs_max
     bnd%x_bs = x_bs
     bnd%sigma_bs = sigma_bs
     bnd%phi_bs = phi_bs
     bnd%i_rescale = i_rescale
     bnd%beta_hartree = beta_hartree
     bnd%delta_inlet = delta_inlet
     bnd%t_resc = t_resc
     call bl_check(bl_set_boundary(h, bnd), 'bl_set_boundary')
  end if
  ! the wall model (compute_wall_model wallmodel.f90:25-144), computed on the
  ! device at every RK stage
  if (iwall_model /= 0 .or. istress_model /= 0) then
     wmp%iwall_model = iwall_model
     wmp%istress_model = istress_model
     wmp%alpha_mean_x = alphas(1)
     wmp%alpha_mean_y = alphas(2)
     wmp%alpha_mean_z = alphas(3)
     wmp%alpha_std = alphas(4)
     wmp%freq_mult = alphas(5)
     wmp%iwall_model_nut = iwall_model_nut
     wmp%frac_vis = frac_vis
     call bl_check(bl_set_wall_model(h, wmp), 'bl_set_wall_model')
  end if
  write(*,*) 'grid ', nx, ny, nz, ' RKscheme ', itime_step, ' LES ', les_model
  call system_clock(c0, crate)
  time1 = real(c0, real64) / real(crate, real64)

  ! time loop (main.f90:84-113).  Nothing is downloaded per step: the
  ! statistics are reduced on the device, and a snapshot's copy to the host
  ! runs under the next step's kernels (written once that step is queued).
  snap_pending = .false.
  snap_step = 0
  snap_t = 0d0
  do istep = 1, nsteps
     call bl_check(bl_compute_dt(h, dt), 'bl_compute_dt')
     call bl_check(bl_step(h), 'bl_step')
     if (snap_pending) then
        call write_snapshot(h, snap_step, snap_t)
        snap_pending = .false.
     end if
     call bl_check(bl_get_time(h, t_dev), 'bl_get_time')
     t = t_dev
     if (mod(istep, nstats) == 0 .or. istep == 1) then
        call compute_statistics(h)
        call output_statistics()
     end if
     if (mod(istep, nmonitor) == 0 .or. istep == 1) then
        call bl_check(bl_monitor(h, mon), 'bl_monitor')
        call bl_check(bl_max_divergence(h, div), 'bl_max_divergence')
        call system_clock(c1)
        time2 = real(c1, real64) / real(crate, real64)
        write(*,*) 'step number  :', istep
        write(*,*) 'time         :', t
        write(*,*) 'time step    :', dt
        write(*,*) ' '
        write(*,*) 'Mean Cf:     :', sum(Cf) / real(nx, real64)
        write(*,*) ' '
        write(*,*) 'Maximum U    :', mon(5)
        write(*,*) 'Maximum V    :', mon(6)
        write(*,*) 'Maximum W    :', mon(7)
        write(*,*) 'Mean U       :', mon(1) / real(int(nxm, int64) * nym * nzm, real64)
        write(*,*) 'Mean V       :', mon(2) / real(int(nxm, int64) * nym * nzm, real64)
        write(*,*) 'Mean W       :', mon(3) / real(int(nxm, int64) * nym * nzm, real64)
        if (les_model > 0) then
           write(*,*) ' '
           write(*,*) 'Maximum nu_t : ', mon(8)
           write(*,*) 'Average nu_t : ', mon(4) / real(int(nxm, int64) * nym * nzm, real64)
        end if
        write(*,*) ' '
        write(*,*) 'Inflow Reynolds numbers:'
        write(*,*) '                Re_x        : ', Rex_inlet
        write(*,*) '                Re_theta    : ', Retheta_inlet
        write(*,*) '                Re_delta    : ', Redelta_inlet
        write(*,*) '                delta99_i   : ', delta99_inlet
        write(*,*) ' '
        write(*,*) 'Maximum divergence          :', div
        write(*,*) 'Elapsed time (s)            :', time2 - time1
        write(*,*) '------------------------------------------------------'
        time1 = time2
     end if
     if (mod(istep, nsave) == 0) then
        call bl_check(bl_snapshot_begin(h), 'bl_snapshot_begin')
        snap_pending = .true.
        snap_step = istep
        snap_t = t
     end if
  end do
  if (snap_pending) call write_snapshot(h, snap_step, snap_t)

  call bl_destroy(h)
  i = 0
end program bl_driver
