! mc_gpu.f90 -- driver of the MI355X engine with the reference's command line, input file and
! output layout:  mc_gpu.x seedname  reads seedname.mcin and writes seedname.mcout,
! seedname_mc.dat and seedname_sconfig.dat.
!
! Plain MC follows mcarlo (mcarlo.F90:463-859) with every temperature an independent chain on
! the device (the reference carries the configuration from one temperature to the next on a
! rank; after the warm-up both sample the same distribution).  Parallel tempering follows
! pt_mcarlo (pt.F90:504-1134): warmup1 plain, warmup2 with exchanges, main phase with
! exchanges and measurements; exchanges swap temperature labels on the device.
program mc_gpu
  use iso_c_binding
  use espins_hip
  use mc_input
  use mc_rng
  implicit none

  real(dp), parameter :: pi = 3.141592653589793238462643383279_dp   ! constants.F90:30
  real(dp), parameter :: twopi = 2 * pi

  type(input_t) :: inp
  type(c_ptr) :: ctx
  character(len=256) :: seedname, arg
  integer :: narg, k, i, a, N, R, out_u, mc_u, sc_u, rc, ncell
  integer(c_int) :: zJ, zB, zD, info(6), mode
  logical :: learned
  integer(c_int), allocatable, target :: jc(:), jn(:), bc(:), bn(:), dc(:), dn(:)
  real(c_double), allocatable, target :: jv(:), bv(:), dvv(:), mag(:), si_par(:), si_ax(:,:), h_par(:), h_ax(:,:)
  real(c_double), allocatable :: spins(:,:,:), tems(:), acc(:,:), e0(:), edir(:), erun(:)
  integer(c_int64_t), allocatable :: rej_w(:), rej(:), nswap(:,:)
  integer(c_int), allocatable :: smap(:)
  integer(c_int64_t) :: nmeas, seed, sweep
  integer :: w1, nsw_warm, nsw, steps_mc
  integer(8) :: t0, t1, rate
  real(dp) :: beta, C, chi, UE, UM, dE, dM, nrm, r1, r2
  type(mt_t) :: rng
  integer :: rank_seed
  real(c_double), allocatable :: e_init(:), e_slot(:), d_slot(:)

  narg = command_argument_count()
  if (narg < 1) then
    write(*, '(a)') 'usage: mc_gpu.x seedname'
    stop 1
  end if
  call get_command_argument(narg, arg)
  seedname = arg
  k = index(seedname, '.mcin', back=.true.)
  if (k > 0) seedname = seedname(1:k-1)
  call system_clock(t0, rate)

  call read_input(trim(seedname)//'.mcin', inp)
  open(newunit=out_u, file=trim(seedname)//'.mcout', status='replace', action='write')
  write(out_u, '(1x,a)') 'ESpinS-compatible Monte Carlo on MI355X (libespins_hip)'
  write(out_u, '(1x,a,i6)') 'Library version            : ', esp_version()
  write(out_u, '(1x,a,3i4)') 'Supercell size             : ', inp%L
  write(out_u, '(1x,a,i6)') 'Atoms in the unit cell     : ', inp%nb
  write(out_u, '(1x,a,l2)') 'Parallel tempering         : ', inp%pt
  write(out_u, '(1x,a,a)')  'Monte Carlo mode           : ', trim(inp%mcarlo_mode)
  write(out_u, '(1x,a,a)')  'Initial spins              : ', trim(inp%initial)

  ! ---- tables (jij_get / bij_get / dij_get through esp_build_table) -------------------
  ncell = inp%L(1) * inp%L(2) * inp%L(3)
  N = ncell * inp%nb
  allocate(mag(N))
  do i = 1, N
    mag(i) = inp%moment(mod(i - 1, inp%nb) + 1)
  end do
  call build(inp%J, 1, zJ, jc, jn, jv)
  zB = 0
  zD = 0
  if (inp%have_B) call build(inp%B, 1, zB, bc, bn, bv)
  if (inp%have_D) call build(inp%D, 3, zD, dc, dn, dvv)
  if (inp%have_si) then
    allocate(si_par(inp%nb), si_ax(3, inp%nb))
    si_par = inp%si_par / k_B                       ! parameters.F90:958
    si_ax = inp%si_axis
  end if
  if (inp%have_h) then
    allocate(h_par(inp%nb), h_ax(3, inp%nb))
    h_par = inp%h_par * bohr_magn / k_B             ! parameters.F90:994
    h_ax = inp%h_axis
  end if

  call check(esp_create(ctx, 0_c_int), 'esp_create')
  call check(esp_set_supercell(ctx, inp%L), 'esp_set_supercell')
  call check(esp_set_model(ctx, int(N, c_int), int(inp%nb, c_int), mag, zJ, jc, jn, jv, &
             zB, optp_i(bc, zB > 0), optp_i(bn, zB > 0), optp_d(bv, zB > 0), &
             zD, optp_i(dc, zD > 0), optp_i(dn, zD > 0), optp_d(dvv, zD > 0), &
             merge(1_c_int, 0_c_int, inp%have_si), optp_d(si_par, inp%have_si), optp_d2(si_ax, inp%have_si), &
             merge(1_c_int, 0_c_int, inp%have_h), optp_d(h_par, inp%have_h), optp_d2(h_ax, inp%have_h)), &
             'esp_set_model')
  call check(esp_get_model_info(ctx, info), 'esp_get_model_info')
  write(out_u, '(1x,a,i10)') 'Spins per replica          : ', N
  write(out_u, '(1x,a,i6,a,i4,a,l2)') 'Bond width', info(2), '  colours', info(3), '  LDS-resident', info(5) /= 0

  ! ---- replicas and initial configuration (mcarlo.F90:268-322, pt.F90:440-491) ----------
  R = inp%tems_num
  allocate(tems(R), spins(3, N, R), acc(6, R), e0(R), edir(R), erun(R), rej_w(R), rej(R), &
           nswap(2, R), smap(R))
  tems = inp%tems
  ! the rank's random seed: the seeds keyword, else /dev/urandom (comms_init_random,
  ! comms.F90:979-1007); it seeds the host MT19937 stream and the device Philox key
  rank_seed = inp%seed
  if (.not. inp%have_seed) rank_seed = urandom_seed()
  seed = int(rank_seed, c_int64_t)
  write(out_u, '(1x,a,i12)') 'Random seed                : ', rank_seed
  call mt_seed(rng, rank_seed)
  if (inp%initial(1:5) == 'ferro') then
    spins = 0
    do k = 1, R
      spins(3, :, k) = mag
    end do
  else if (inp%initial(1:4) == 'rand') then
    ! the reference's 'rand' configuration (mcarlo.F90:275-286, pt.F90:448-459): two draws of
    ! the rank's MT19937 stream per site in site order, not uniform on the sphere, the same for
    ! every replica
    do i = 1, N
      r1 = mt_real2(rng)
      r2 = mt_real2(rng)
      spins(1, i, 1) = mag(i) * sin(r1 * pi) * cos(r2 * twopi)
      spins(2, i, 1) = mag(i) * sin(r1 * pi) * sin(r2 * twopi)
      spins(3, i, 1) = mag(i) * cos(r1 * pi)
    end do
    do k = 2, R
      spins(:, :, k) = spins(:, :, 1)
    end do
  else
    call read_sconfig(trim(seedname)//'_sconfig.dat')
  end if
  call check(esp_set_replicas(ctx, int(R, c_int), tems, spins), 'esp_set_replicas')
  call check(esp_energy(ctx, e0), 'esp_energy')
  mode = merge(ESP_MODE_CONST, ESP_MODE_RANDOM, inp%mcarlo_mode(1:5) == 'const')
  learned = inp%mcarlo_mode(1:5) == 'learn'   ! warm-up with 'random', then train, then learned

  if (.not. inp%pt) then
    ! ---- plain MC: warm-up then measured steps, all temperatures at once ----------------
    sweep = 0
    if (inp%steps_warmup > 0) then
      call check(esp_run(ctx, 1_c_int, int(inp%steps_warmup, c_int), int(inp%steps_measure, c_int), &
                 0_c_int, 0_c_int, 1_c_int, mode, real(inp%tilt, c_double), seed, sweep, 0_c_int64_t, &
                 0_c_int), 'esp_run warmup')
      sweep = sweep + int(inp%steps_warmup, c_int64_t) * inp%steps_measure
    end if
    if (learned) call train_proposal()
    call check(esp_get_stats(ctx, acc, rej_w, erun, nmeas), 'esp_get_stats')
    call check(esp_reset_stats(ctx), 'esp_reset_stats')
    if (inp%energy_write) call check(esp_set_trace(ctx, int(inp%steps_mc, c_int)), 'esp_set_trace')
    call check(esp_run(ctx, 1_c_int, int(inp%steps_mc, c_int), int(inp%steps_measure, c_int), &
               1_c_int, 0_c_int, 1_c_int, mode, real(inp%tilt, c_double), seed, sweep, 0_c_int64_t, &
               0_c_int), 'esp_run measure')
    steps_mc = inp%steps_mc
    if (inp%energy_write) call write_energy_series(steps_mc)
    call check(esp_get_stats(ctx, acc, rej, erun, nmeas), 'esp_get_stats')
    call check(esp_energy(ctx, edir), 'esp_energy')
    call check(esp_get_spins(ctx, 0_c_int, spins), 'esp_get_spins')
    call write_mc_tables(int(inp%steps_warmup, 8) * inp%steps_measure * N, &
                         int(steps_mc, 8) * inp%steps_measure * N, .false.)
  else
    ! ---- parallel tempering schedule (pt.F90:529-544) ------------------------------------
    allocate(e_init(R), e_slot(R), d_slot(R))
    w1 = ceiling(real(inp%steps_warmup, dp) / 2.0_dp)
    nsw_warm = ceiling(real(w1, dp) / real(inp%pt_steps_swap, dp))
    if (inp%pt_steps_swap * nsw_warm /= w1) then
      write(out_u, '(1x,a)') '***** The number of Warm-up steps is not diviable to pt_steps_swap : '
      write(out_u, '(1x,a,I9)') '***** For consistancy the number of Warm-up steps increases to     : ', &
                                2 * nsw_warm * inp%pt_steps_swap
    end if
    w1 = nsw_warm * inp%pt_steps_swap
    nsw = ceiling(real(inp%steps_mc, dp) / real(inp%pt_steps_swap, dp))
    if (nsw * inp%pt_steps_swap /= inp%steps_mc) then
      write(out_u, '(1x,a)') '***** The number of Monte Carlo steps is not diviable to pt_steps_swap : '
      write(out_u, '(1x,a,I9)') '***** For consistancy the number of Monte Carlo steps increases to     : ', &
                                nsw * inp%pt_steps_swap
    end if
    steps_mc = nsw * inp%pt_steps_swap
    sweep = 0
    if (w1 > 0) then            ! warmup1: plain sweeps (pt.F90:557-629)
      call check(esp_run(ctx, 1_c_int, int(w1, c_int), int(inp%steps_measure, c_int), 0_c_int, 0_c_int, &
                 1_c_int, mode, real(inp%tilt, c_double), seed, sweep, 0_c_int64_t, 0_c_int), 'warmup1')
      sweep = sweep + int(w1, c_int64_t) * inp%steps_measure
      call check(esp_get_stats(ctx, acc, rej_w, erun, nmeas), 'esp_get_stats')
      call check(esp_energy(ctx, edir), 'esp_energy')
      write(out_u, '(/1x,a)') '*------------------------------------ WARMUP1 --------------------------------------*'
      write(out_u, '(1x,a)') '|     Temp      Total Steps    Rejected Steps    Accepted Steps   Acc/Tot Ratio     |'
      write(out_u, '(1x,a)') '+-----------------------------------------------------------------------------------+'
      do k = 1, R               ! no exchanges yet: replica k is at temperature k
        write(out_u, 224) tems(k), int(w1, 8) * inp%steps_measure * N, rej_w(k), &
                          int(w1, 8) * inp%steps_measure * N - rej_w(k), &
                          real(int(w1, 8) * inp%steps_measure * N - rej_w(k), dp) / &
                          real(int(w1, 8) * inp%steps_measure * N, dp), '  <-- WARMUP1'
      end do
      call energy_table('WARMUP1', e0, erun, edir, 'ENERGY1')
      call check(esp_get_spins(ctx, 1_c_int, spins), 'esp_get_spins')
      call write_sconfig(trim(seedname)//'_sconfig_warmup.dat', .false.)
      ! warmup2: exchanges every pt_steps_swap MC steps, parity 1,2,1,... (pt.F90:654-713)
      call check(esp_reset_stats(ctx), 'esp_reset_stats')
      e_init = erun
      write(out_u, '(/1x,a)') '*---------------------------------- SWAP in WARMUP2 --------------------------------*'
      write(out_u, '(1x,a)') '|     Num_Steps                     Configuration Numbers                           |'
      write(out_u, '(1x,a)') '+-----------------------------------------------------------------------------------+'
      call check(esp_get_map(ctx, smap, nswap), 'esp_get_map')
      write(out_u, *) 'Initial configuration = ', smap
      call pt_rounds(nsw_warm, w1, 0_c_int, 0, 'warmup2')
      write(out_u, *) 'Final configuration = ', smap
      write(out_u, '(1x,85a)') '*-----------------------------------------------------------------------------------*'
      call check(esp_get_stats(ctx, acc, rej_w, erun, nmeas), 'esp_get_stats')
      call check(esp_get_map(ctx, smap, nswap), 'esp_get_map')
      call check(esp_energy(ctx, edir), 'esp_energy')
      call swap_table('WARMUP2')
      write(out_u, '(/1x,a)') '*------------------------------------ WARMUP2 --------------------------------------*'
      write(out_u, '(1x,a)') '|     Temp      Total Steps    Rejected Steps    Accepted Steps   Acc/Tot Ratio     |'
      write(out_u, '(1x,a)') '+-----------------------------------------------------------------------------------+'
      do k = 1, R               ! rejections are counted by temperature slot
        write(out_u, 224) tems(k), int(w1, 8) * inp%steps_measure * N, rej_w(k), &
                          int(w1, 8) * inp%steps_measure * N - rej_w(k), &
                          real(int(w1, 8) * inp%steps_measure * N - rej_w(k), dp) / &
                          real(int(w1, 8) * inp%steps_measure * N, dp), '  <-- WARMUP2'
      end do
      ! energies follow the configurations (by slot now); the initial column stays by slot
      do k = 1, R
        e_slot(k) = erun(smap(k))
        d_slot(k) = edir(smap(k))
      end do
      call energy_table('WARMUP2', e_init, e_slot, d_slot, 'ENERGY2')
      sweep = sweep + int(w1, c_int64_t) * inp%steps_measure
    else
      write(out_u, '(/1x,85a)') '+------------------------------------ WARMUP --------------------------------------+'
      write(out_u, '(1x,85a)') '|                                                                                   |'
      write(out_u, '(1x,85a)') '+-----------------------------------------------------------------------------------+'
      write(out_u, '(1x,85a)') ' No Warm-up step is done! '
      write(out_u, '(1x,a)') '*--------------------------------- END of WARMUP -----------------------------------*'
    end if
    if (learned) call train_proposal()
    call check(esp_get_stats(ctx, acc, rej_w, erun, nmeas), 'esp_get_stats')
    call check(esp_get_map(ctx, smap, nswap), 'esp_get_map')
    do k = 1, R
      e_init(k) = erun(smap(k))   ! running energy of the configuration at each slot
    end do
    call check(esp_reset_stats(ctx), 'esp_reset_stats')
    if (inp%energy_write) call check(esp_set_trace(ctx, int(steps_mc, c_int)), 'esp_set_trace')
    ! main phase: measurement after every MC step, exchange every pt_steps_swap steps,
    ! parity restarts at 1 (pt.F90:773-908)
    write(out_u, '(/1x,85a)') '*------------------------------ SWAP in Monte-Carlo --------------------------------*'
    write(out_u, '(1x,85a)') '|     Num_Steps                     Configuration Numbers                           |'
    write(out_u, '(1x,85a)') '+-----------------------------------------------------------------------------------+'
    write(out_u, *) 'Initial configuration = ', smap
    call pt_rounds(nsw, steps_mc, 1_c_int, nsw_warm, 'main')
    write(out_u, *) 'Final configuration = ', smap
    write(out_u, '(1x,85a)') '*-----------------------------------------------------------------------------------*'
    if (inp%energy_write) call write_energy_series(steps_mc)
    call check(esp_get_stats(ctx, acc, rej, erun, nmeas), 'esp_get_stats')
    call check(esp_get_map(ctx, smap, nswap), 'esp_get_map')
    call check(esp_energy(ctx, edir), 'esp_energy')
    call check(esp_get_spins(ctx, 1_c_int, spins), 'esp_get_spins')
    write(out_u, *) ''
    call swap_table('MONTECARLO')
    call write_mc_tables(0_8, int(steps_mc, 8) * inp%steps_measure * N, .true.)
  end if

  ! ---- _sconfig.dat (mcarlo.F90:109-113, 813-816) ----------------------------------------
  call write_sconfig(trim(seedname)//'_sconfig.dat', .true.)

  call check(esp_destroy(ctx), 'esp_destroy')
  call system_clock(t1)
  write(out_u, '(/1x,a,f12.3,a)') 'Time for montecarlo ', real(t1 - t0, dp) / real(rate, dp), ' (sec)'
  close(out_u)
224 format(3x,f9.4,1x,i14,1x,i14,5x,i14,7x,f10.8,4x,a)   ! pt.F90:1126

contains

  ! seedname_energy-TNNN.dat, one file per temperature (open_E_unit / write_E / close_E_unit,
  ! get_quant.F90:1130-1223): header, then the running energy per site of every measured MC
  ! step, in temperature order.
  subroutine write_energy_series(nsteps)
    integer, intent(in) :: nsteps
    real(c_double), allocatable :: tr(:,:,:)
    integer(c_int) :: nrec
    integer :: t, s, u
    character(len=8) :: d
    character(len=10) :: tm
    character(len=300) :: fname
    allocate(tr(2, R, nsteps))
    call check(esp_get_trace(ctx, tr, nrec), 'esp_get_trace')
    call date_and_time(date=d, time=tm)
    do t = 1, R
      write(fname, '(a,a,i3.3,a)') trim(seedname), '_energy-T', t, '.dat'
      open(newunit=u, file=trim(fname), status='replace', action='write')
      write(u, *) '#written on '//d(7:8)//'/'//d(5:6)//'/'//d(1:4)//' at '//tm(1:2)//':'//tm(3:4)//':'//tm(5:6)
      write(u, '(3x,a,f9.4)') '# T = ', tems(t)
      write(u, '(3x,i16)') nsteps
      do s = 1, min(int(nrec), nsteps)
        write(u, '(I14,4x,1E16.8)') s, tr(1, t, s) / real(N, dp)
      end do
      close(u)
    end do
    call check(esp_set_trace(ctx, 0_c_int), 'esp_set_trace')
    deallocate(tr)
  end subroutine

  ! Learned proposal: records of local conditionals from the warmed-up chains (plain sweeps
  ! between collections), Adam training on the device, then the measured phase runs with
  ! mode = ESP_MODE_LEARNED.  Initial weights: He-normal hidden layers, a small output layer.
  subroutine train_proposal()
    real(c_float) :: w(ESP_LEARN_NPARAMS)
    real(c_double) :: losses(1)
    real(dp) :: u1, u2, g, sd
    integer :: p, e, nrec, per, got, sz(8), nseed
    integer, allocatable :: rs(:)
    integer(c_int64_t) :: s0
    ! layer sizes: W1 32x5, b1 32, W2 32x32, b2 32, W3 32x32, b3 32, W4 6x32, b4 6
    sz = [160, 32, 1024, 32, 1024, 32, 192, 6]
    call random_seed(size=nseed)
    allocate(rs(nseed))
    rs = int(seed) + 17 * [(p, p = 1, nseed)]
    call random_seed(put=rs)
    p = 0
    do e = 1, 8
      sd = 0.0_dp
      if (e == 1) sd = sqrt(2.0_dp / 5.0_dp)
      if (e == 3 .or. e == 5) sd = sqrt(2.0_dp / 32.0_dp)
      if (e == 7) sd = 0.01_dp
      do k = 1, sz(e)
        call random_number(u1)
        call random_number(u2)
        g = sqrt(-2.0_dp * log(1.0_dp - u1)) * cos(2.0_dp * acos(-1.0_dp) * u2)
        p = p + 1
        w(p) = real(sd * g, c_float)
      end do
    end do
    w(ESP_LEARN_NPARAMS - 5:ESP_LEARN_NPARAMS) = [0.0, 0.5, 0.0, -0.5, -0.5, 0.5]
    call check(esp_learn_set_weights(ctx, w, ESP_LEARN_NPARAMS), 'esp_learn_set_weights')
    call check(esp_learn_clear(ctx), 'esp_learn_clear')
    nrec = inp%learned_records
    per = max(1, nrec / 16)
    got = 0
    s0 = seed * 1048576_c_int64_t
    do while (got < nrec)
      call check(esp_learn_collect(ctx, int(min(per, nrec - got), c_int), s0 + got), 'esp_learn_collect')
      got = got + min(per, nrec - got)
      call check(esp_sweep(ctx, 2_c_int, ESP_MODE_RANDOM, real(inp%tilt, c_double), seed, sweep), 'esp_sweep')
      sweep = sweep + 2
    end do
    write(out_u, '(/1x,a,i10,a,i4,a)') 'Learned proposal: training on ', nrec, ' records, ', &
      inp%learned_epochs, ' epochs'
    do e = 1, inp%learned_epochs
      call check(esp_learn_train(ctx, 1_c_int, 4096_c_int, real(2.0e-3_dp * 0.5_dp**(e - 1), c_float), &
                 seed + e, losses), 'esp_learn_train')
      write(out_u, '(1x,a,i4,a,f14.6)') '   epoch ', e, '  mean -log q = ', losses(1)
    end do
    mode = ESP_MODE_LEARNED
  end subroutine

  ! A seed from /dev/urandom when the input gives none (comms_init_random, comms.F90:995-999).
  integer function urandom_seed() result(v)
    integer :: u, ios
    integer(c_int64_t) :: c
    v = 0
    open(newunit=u, file='/dev/urandom', access='stream', form='unformatted', action='read', iostat=ios)
    if (ios == 0) then
      read(u, iostat=ios) v
      close(u)
    end if
    if (ios /= 0) then
      call system_clock(c)
      v = int(mod(c, 2147483647_c_int64_t))
    end if
    v = abs(v)   ! the reference's MT stream equals standard MT19937 for non-negative seeds
  end function

  ! _sconfig.dat / _sconfig_warmup.dat: date line, (supercell line), N and the temperature
  ! count, then per temperature T (F12.6) and the spins (3F12.6) (mcarlo.F90:109-114,813-816;
  ! pt.F90:610-629).  spins must be by temperature slot.
  subroutine write_sconfig(fname, with_cell)
    character(len=*), intent(in) :: fname
    logical, intent(in) :: with_cell
    character(len=8) :: d
    character(len=10) :: tm
    call date_and_time(date=d, time=tm)
    open(newunit=sc_u, file=fname, status='replace', action='write')
    write(sc_u, *) 'Created on '//d(7:8)//'/'//d(5:6)//'/'//d(1:4)//' at '//tm(1:2)//':'//tm(3:4)//':'//tm(5:6)
    if (with_cell) write(sc_u, '(3I6)') inp%L
    write(sc_u, '(2I6)') N, R
    do k = 1, R
      write(sc_u, '(F12.6)') tems(k)
      do i = 1, N
        write(sc_u, '(3F12.6)') spins(:, i, k)
      end do
    end do
    close(sc_u)
  end subroutine

  ! PT energy table (format 127): T, initial, final by adding, final direct, by slot.
  subroutine energy_table(phase, ei, ea, ed, tag)
    character(len=*), intent(in) :: phase, tag
    real(c_double), intent(in) :: ei(:), ea(:), ed(:)
    integer :: t
    write(out_u, '(/1x,a)') '*------------------------------------ '//phase//' --------------------------------------*'
    write(out_u, '(1x,a)') '|   Temp             E_init             E_final(by adding)       E_final(direct)    |'
    write(out_u, '(1x,a)') '+-----------------------------------------------------------------------------------+'
    do t = 1, R
      write(out_u, 227) tems(t), ei(t), ea(t), ed(t), '  <-- '//tag
    end do
    write(out_u, '(1x,a)') '*---------------------------------- END of '//phase//' ---------------------------------*'
227 format(2x,f9.4,8x,E15.8,8x,E15.8,8x,E15.8,2x,a)
  end subroutine

  ! nrounds exchange rounds (pt_steps_swap MC steps each) of the phase that ends after
  ! nsteps MC steps, printing the replica map at the start of MC step 1, of every multiple of
  ! pt_print_swap and of the last step (pt.F90:661-662,780-781).  Rounds between prints run
  ! as one graph-replayed esp_run.
  subroutine pt_rounds(nrounds, nsteps, measure, step0, what)
    integer, intent(in) :: nrounds, nsteps, step0
    integer(c_int), intent(in) :: measure
    character(len=*), intent(in) :: what
    integer :: done, n, kr, target
    integer(8) :: p
    done = 0
    n = 1
    do while (n <= nsteps)
      kr = (n - 1) / inp%pt_steps_swap        ! exchange rounds before MC step n
      if (kr > done) then
        call check(esp_run(ctx, int(kr - done, c_int), int(inp%pt_steps_swap, c_int), &
                   int(inp%steps_measure, c_int), measure, 1_c_int, int(1 + mod(done, 2), c_int), mode, &
                   real(inp%tilt, c_double), seed, sweep + int(done, c_int64_t) * inp%pt_steps_swap * &
                   inp%steps_measure, int(step0 + done, c_int64_t), 1_c_int), what)
        done = kr
        call check(esp_get_map(ctx, smap, nswap), 'esp_get_map')
      end if
      write(out_u, *) n, ' | ', smap
      ! next printed step: the next multiple of pt_print_swap, or the last step
      p = (int(n, 8) / max(1, inp%pt_print_swap) + 1) * max(1, inp%pt_print_swap)
      if (n == nsteps) exit
      target = int(min(p, int(nsteps, 8)))
      n = target
    end do
    if (nrounds > done) then
      call check(esp_run(ctx, int(nrounds - done, c_int), int(inp%pt_steps_swap, c_int), &
                 int(inp%steps_measure, c_int), measure, 1_c_int, int(1 + mod(done, 2), c_int), mode, &
                 real(inp%tilt, c_double), seed, sweep + int(done, c_int64_t) * inp%pt_steps_swap * &
                 inp%steps_measure, int(step0 + done, c_int64_t), 1_c_int), what)
      call check(esp_get_map(ctx, smap, nswap), 'esp_get_map')
    end if
  end subroutine

  subroutine check(st, what)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: what
    if (st /= ESP_OK) then
      write(*, '(1x,a,i4,2a)') 'Error: '//what//' failed with status ', st, ': ', trim(esp_error_text(ctx))
      error stop 2
    end if
  end subroutine

  ! jij_get: rows (eV) -> ELL tables in Kelvin (jij.F90:163)
  subroutine build(r, ncomp, z, cnt, nbr, val)
    type(rows_t), intent(in) :: r
    integer, intent(in) :: ncomp
    integer(c_int), intent(out) :: z
    integer(c_int), allocatable, intent(out) :: cnt(:), nbr(:)
    real(c_double), allocatable, intent(out) :: val(:)
    real(c_double), allocatable, target :: dvec(:,:)
    real(c_double), allocatable :: v(:)
    integer(c_int) :: zcap, st
    type(c_ptr) :: pd
    zcap = max(r%n, 1)
    allocate(cnt(N), nbr(N * zcap), val(N * zcap * ncomp), v(max(r%n, 1)))
    v = 0
    if (r%n > 0) v(1:r%n) = r%val / k_B
    pd = c_null_ptr
    if (ncomp == 3) then
      allocate(dvec(3, max(r%n, 1)))
      dvec = 0
      if (r%n > 0) dvec(:, 1:r%n) = r%dvec
      pd = c_loc(dvec)
    end if
    if (r%n == 0) then
      z = 0
      cnt = 0
      return
    end if
    st = esp_build_table(int(inp%nb, c_int), inp%basis, inp%moment, inp%L, &
                         merge(1_c_int, 0_c_int, inp%open_boundary), int(r%n, c_int), r%f1, r%f2, v, &
                         pd, int(ncomp, c_int), zcap, cnt, nbr, val, z)
    if (st /= ESP_OK) call input_error('Unrecognised Atomic Position found in a coupling block')
  end subroutine

  function optp_i(x, use) result(p)
    integer(c_int), allocatable, target, intent(in) :: x(:)
    logical, intent(in) :: use
    type(c_ptr) :: p
    p = c_null_ptr
    if (use) p = c_loc(x)
  end function
  function optp_d(x, use) result(p)
    real(c_double), allocatable, target, intent(in) :: x(:)
    logical, intent(in) :: use
    type(c_ptr) :: p
    p = c_null_ptr
    if (use) p = c_loc(x)
  end function
  function optp_d2(x, use) result(p)
    real(c_double), allocatable, target, intent(in) :: x(:,:)
    logical, intent(in) :: use
    type(c_ptr) :: p
    p = c_null_ptr
    if (use) p = c_loc(x)
  end function

  ! mcarlo_read_sconfig (get_quant.F90:543-598): normalise, then scale by the moment
  subroutine read_sconfig(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios, nw, nt, kk, j
    real(dp) :: tt
    character(len=256) :: line
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) call input_error('Initial_sconfig requested from file but '//fname//' not found')
    read(u, '(a)') line
    read(u, '(a)') line
    read(u, *) nw, nt
    if (nw /= N) call input_error('Wrong matrix size in read_sconfig')
    if (nt /= R) call input_error('Wrong temperature number in read_sconfig')
    do kk = 1, nt
      read(u, *) tt
      if (abs(tt - tems(kk)) > 1e-4_dp) call input_error('Mismatch in temperatures: read_sconfig')
      do j = 1, N
        read(u, *) spins(:, j, kk)
        if (all(abs(spins(:, j, kk)) < 1e-4_dp)) call input_error('zero spin in '//fname)
        spins(:, j, kk) = spins(:, j, kk) / sqrt(dot_product(spins(:, j, kk), spins(:, j, kk))) * mag(j)
      end do
    end do
    close(u)
  end subroutine

  subroutine swap_table(phase)
    character(len=*), intent(in) :: phase
    integer :: t
    write(out_u, '(/1x,a)') '*-------------------------------- '//phase//' swaps ---------------------------------*'
    write(out_u, '(1x,a)') '|  Temp_i     Temp_f         Tot-Swaps       Rej-Swaps      Acc-Swaps  Acc/Tot Ratio|'
    do t = 1, R
      write(out_u, 128) tems(t), tems(smap(t)), nswap(1, t), nswap(1, t) - nswap(2, t), nswap(2, t), &
                        real(nswap(2, t), dp) / max(1.0_dp, real(nswap(1, t), dp))
    end do
128 format(2x,f9.4,2x,f9.4,1x,i14,1x,i14,1x,i14,3x,f14.8,2x)
  end subroutine

  subroutine write_mc_tables(nwarm, nmc, pt)
    integer(8), intent(in) :: nwarm, nmc
    logical, intent(in) :: pt
    integer :: t
    open(newunit=mc_u, file=trim(seedname)//'_mc.dat', status='replace', action='write')
    write(mc_u, '(1x,a)') '#written by mc_gpu.x'
    write(mc_u, '(1x,a)') '#------------------------------------------------- MONTECARLO --------'// &
                          '-----------------------------------------#'
    write(mc_u, '(1x,a)') '# Temp      Magnetization     Energy_ave          C_M              Sus'// &
                          '              U_E              U_M       #'
    write(mc_u, '(1x,a)') '#-----------------------------------------------------------------------'// &
                          '---------------------------------------#'
    write(out_u, '(/1x,a)') '*--------------------------------- MONTECARLO --------------------------------------*'
    write(out_u, '(1x,a)') '|     Temp      Total Steps    Rejected Steps    Accepted Steps   Acc/Tot Ratio     |'
    nrm = 1.0_dp / real(nmeas, dp)
    do t = 1, R
      ! final formulas, mcarlo.F90:713-733
      beta = 1.0_dp / tems(t)
      acc(:, t) = acc(:, t) * nrm
      C = beta * beta * (acc(2, t) - acc(1, t)**2) / real(N, dp)
      chi = beta * (acc(5, t) - acc(4, t)**2) * N
      UE = 1.0_dp - (1.0_dp / 3.0_dp) * (acc(3, t) / acc(2, t)**2)
      UM = 1.0_dp - (1.0_dp / 3.0_dp) * (acc(6, t) / acc(5, t)**2)
      dE = sqrt(max(acc(2, t) - acc(1, t)**2, 0.0_dp) * nrm) / real(N, dp)
      dM = sqrt(max(acc(5, t) - acc(4, t)**2, 0.0_dp) * nrm)
      if (nwarm > 0) write(out_u, 124) tems(t), nwarm, rej_w(t), nwarm - rej_w(t), &
                                       real(nwarm - rej_w(t), dp) / real(nwarm, dp), '  <-- WARMUP'
      if (.not. pt) then
        write(out_u, 124) tems(t), nmc, rej(t), nmc - rej(t), real(nmc - rej(t), dp) / real(nmc, dp), &
                          '  <-- MCARLO'
        write(out_u, '(4x,a,E15.7,a)') ' Initial Energy                           = ', e0(t), '  <-- ENERGY'
        write(out_u, '(4x,a,E15.7,a)') ' Final Energy by adding differences       = ', erun(t), '  <-- ENERGY'
        write(out_u, '(4x,a,E15.7,a)') ' Final Energy by direct calculations      = ', edir(t), '  <-- ENERGY'
      else              ! pt.F90:1019-1026: energies of the configuration at each temperature slot
        write(out_u, 224) tems(t), nmc, rej(t), nmc - rej(t), real(nmc - rej(t), dp) / real(nmc, dp), &
                          '  <-- MCARLO'
        write(out_u, '(4x,a,E15.7,a)') ' Initial Energy                           = ', e_init(t), '  <-- ENERGY3'
        write(out_u, '(4x,a,E15.7,a)') ' Final Energy by adding differences       = ', erun(smap(t)), '  <-- ENERGY3'
        write(out_u, '(4x,a,E15.7,a)') ' Final Energy by direct calculations      = ', edir(smap(t)), '  <-- ENERGY3'
      end if
      write(out_u, '(4x,a,f15.9)') '                             Delta E      = +/- ', dE
      write(out_u, '(4x,a,f15.9)') '                             Delta M      = +/- ', dM
      if (pt) write(out_u, '(1x,a85)') repeat('-', 85)
      write(mc_u, 125) tems(t), acc(4, t), acc(1, t) / real(N, dp), C, chi, UE, UM
    end do
    close(mc_u)
124 format(3x,f9.4,1x,i14,1x,i14,5x,i14,7x,f10.8,1x,a)
224 format(3x,f9.4,1x,i14,1x,i14,5x,i14,7x,f10.8,4x,a)
125 format(f9.4,2x,E15.8,2x,E15.8,2x,E15.8,2x,E15.8,2x,E15.8,2x,E15.8)
  end subroutine

end program mc_gpu
