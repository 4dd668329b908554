! mc_gpu.f90 -- driver of the MI355X engine with the reference's command line, input file and
! output layout:  mc_gpu.x seedname  reads seedname.mcin and writes seedname.mcout,
! seedname_mc.dat and seedname_sconfig.dat.  mc_gpu.x -dryrun seedname only reads the input
! and writes the parameter echo.
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
  use mc_echo
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
  integer(8) :: t0, t1, rate, t_start
  real(dp) :: tm_j = 0, tm_b = 0, tm_d = 0, tm_mc = 0   ! stopwatches of the timing table
  real(dp) :: beta, C, chi, UE, UM, dE, dM, nrm, r1, r2
  type(mt_t) :: rng
  integer :: rank_seed
  real(c_double), allocatable :: e_init(:), e_slot(:), d_slot(:)
  integer(c_int) :: nfreq
  integer :: nshell
  integer :: sg_used = 0                 ! seed of the spin-glass normals
  logical :: dryrun = .false.            ! -dryrun: read the input, write the echo, stop

  narg = command_argument_count()
  if (narg < 1) then
    write(*, '(a)') 'usage: mc_gpu.x [-dryrun] seedname'
    stop 1
  end if
  do k = 1, narg - 1
    call get_command_argument(k, arg)
    if (trim(arg) == '-dryrun') dryrun = .true.
  end do
  call get_command_argument(narg, arg)
  seedname = arg
  k = index(seedname, '.mcin', back=.true.)
  if (k > 0) seedname = seedname(1:k-1)
  call system_clock(t0, rate)
  t_start = t0

  call read_input(trim(seedname)//'.mcin', inp)
  open(newunit=out_u, file=trim(seedname)//'.mcout', status='replace', action='write')
  write(out_u, '(1x,a,i6,a)') 'ESpinS-compatible Monte Carlo on MI355X (libespins_hip version', esp_version(), ')'
  write(out_u, '(/,1x,a)') 'Running on one GPU'
  call write_echo(out_u, inp)
  call system_clock(t1)
  write(out_u, '(1x,a25,f11.3,a)') 'Time to read parameters  ', real(t1 - t0, dp) / real(rate, dp), ' (sec)'
  t0 = t1
  if (dryrun) then
    write(out_u, '(/a)') ' Exiting... '//trim(seedname)//'.mcin read (-dryrun).'
    close(out_u)
    stop
  end if

  ! ---- tables (jij_get / bij_get / dij_get through esp_build_table) -------------------
  ncell = inp%L(1) * inp%L(2) * inp%L(3)
  N = ncell * inp%nb
  allocate(mag(N))
  do i = 1, N
    mag(i) = inp%moment(mod(i - 1, inp%nb) + 1)
  end do
  call build(inp%J, 1, zJ, jc, jn, jv)
  if (inp%spin_glass) call disorder()
  zB = 0
  zD = 0
  call system_clock(t1)
  tm_j = real(t1 - t0, dp) / real(rate, dp)
  write(out_u, '(1x,a25,f11.3,a)') 'Time to get Jij matrix   ', tm_j, ' (sec)'
  t0 = t1
  if (inp%have_B) then
    call build(inp%B, 1, zB, bc, bn, bv)
    call system_clock(t1)
    tm_b = real(t1 - t0, dp) / real(rate, dp)
    write(out_u, '(1x,a25,f11.3,a)') 'Time to get Bij matrix   ', tm_b, ' (sec)'
    t0 = t1
  end if
  if (inp%have_D) then
    call build(inp%D, 3, zD, dc, dn, dvv)
    call system_clock(t1)
    tm_d = real(t1 - t0, dp) / real(rate, dp)
    write(out_u, '(1x,a25,f11.3,a)') 'Time to get Dij matrix   ', tm_d, ' (sec)'
    t0 = t1
  end if
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
  call write_seed_box(out_u, inp, rank_seed, sg_used)
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
  call setup_observables()
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
    if (inp%energy_write .or. inp%lbinerror) &
      call check(esp_set_trace(ctx, int(inp%steps_mc, c_int)), 'esp_set_trace')
    call check(esp_run(ctx, 1_c_int, int(inp%steps_mc, c_int), int(inp%steps_measure, c_int), &
               1_c_int, 0_c_int, 1_c_int, mode, real(inp%tilt, c_double), seed, sweep, 0_c_int64_t, &
               0_c_int), 'esp_run measure')
    steps_mc = inp%steps_mc
    call write_observables(.false., steps_mc)
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
    if (inp%energy_write .or. inp%lbinerror) call check(esp_set_trace(ctx, int(steps_mc, c_int)), 'esp_set_trace')
    ! main phase: measurement after every MC step, exchange every pt_steps_swap steps,
    ! parity restarts at 1 (pt.F90:773-908)
    write(out_u, '(/1x,85a)') '*------------------------------ SWAP in Monte-Carlo --------------------------------*'
    write(out_u, '(1x,85a)') '|     Num_Steps                     Configuration Numbers                           |'
    write(out_u, '(1x,85a)') '+-----------------------------------------------------------------------------------+'
    write(out_u, *) 'Initial configuration = ', smap
    call pt_rounds(nsw, steps_mc, 1_c_int, nsw_warm, 'main')
    write(out_u, *) 'Final configuration = ', smap
    write(out_u, '(1x,85a)') '*-----------------------------------------------------------------------------------*'
    call write_observables(.true., steps_mc)
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
  tm_mc = real(t1 - t0, dp) / real(rate, dp)
  if (inp%pt) then                  ! mc_prog.F90:170-185
    write(out_u, '(/1x,a,f11.3,a)') 'Time for Parallel-Tempering montecarlo ', tm_mc, ' (sec)'
  else
    write(out_u, '(/1x,a,f11.3,a)') 'Time for montecarlo ', tm_mc, ' (sec)'
  end if
  ! closing lines and the stopwatch table (mc_prog.F90:199-208, io_print_timings io.F90:130-146)
  write(out_u, '(/,1x,a25,f11.3,a)') 'Total Execution Time     ', real(t1 - t_start, dp) / real(rate, dp), ' (sec)'
  write(out_u, '(/1x,a)') '*===================================================================================*'
  write(out_u, '(1x,a)')  '|                                 TIMING INFORMATION                                |'
  write(out_u, '(1x,a)')  '*===================================================================================*'
  write(out_u, '(1x,a)')  '|    Tag                                                Ncalls      Time (s)        |'
  write(out_u, '(1x,a)')  '|-----------------------------------------------------------------------------------|'
  call timing_row('jij: get', tm_j)
  if (inp%have_B) call timing_row('bij: get', tm_b)
  if (inp%have_D) call timing_row('dij: get', tm_d)
  call timing_row(merge('pt_main    ', 'mcarlo_main', inp%pt), tm_mc)
  write(out_u, '(1x,a)')  '*-----------------------------------------------------------------------------------*'
  if (inp%pt) then
    write(out_u, '(/,1x,a)') 'All done: ptmc exiting'
  else
    write(out_u, '(/,1x,a)') 'All done: mcarlo exiting'
  end if
  close(out_u)
224 format(3x,f9.4,1x,i14,1x,i14,5x,i14,7x,f10.8,4x,a)   ! pt.F90:1126

contains

  subroutine timing_row(tag, secs)
    character(len=*), intent(in) :: tag
    real(dp), intent(in) :: secs
    character(len=48) :: t
    t = tag
    write(out_u, '(1x,"|",a48,":",i10,4x,f10.3,10x,"|")') t, 1, secs
  end subroutine

  ! Spin glass (jij_spin_glass, jij.F90:506-556): the sigma of every slot of the J table from
  ! the sig= of the row that made it (a table of row numbers built over the same rows, zero-J
  ! rows dropped as in the J table), sigma / k_B as jij.F90:164 leaves it, then Gaussian
  ! disorder from the MT19937 stream of spin_glass_seed (/dev/urandom without it).
  subroutine disorder()
    integer(c_int), allocatable :: ic(:), in(:)
    real(c_double), allocatable :: iv(:), ids(:), sg(:), ones(:)
    integer(c_int) :: z, st, bad(2), sgs
    integer :: k
    allocate(ids(inp%J%n), ones(inp%nb))
    ones = 1
    do k = 1, inp%J%n
      ids(k) = merge(real(k, dp), 0.0_dp, inp%J%val(k) /= 0)
    end do
    allocate(ic(N), in(N * max(inp%J%n, 1)), iv(N * max(inp%J%n, 1)))
    st = esp_build_table(int(inp%nb, c_int), inp%basis, ones, inp%L, merge(1_c_int, 0_c_int, inp%open_boundary), &
                         int(inp%J%n, c_int), inp%J%f1, inp%J%f2, ids, c_null_ptr, 1_c_int, &
                         int(max(inp%J%n, 1), c_int), ic, in, iv, z)
    if (st /= ESP_OK .or. z /= zJ) call input_error('spin glass: row table does not match the J table')
    allocate(sg(N * max(zJ, 1)))
    sg = 0
    do k = 1, N * zJ
      if (nint(iv(k)) > 0) sg(k) = inp%J%sigma(nint(iv(k))) / k_B
    end do
    sgs = inp%sg_seed
    if (.not. inp%have_sg_seed) sgs = urandom_seed()
    sg_used = sgs
    st = esp_spin_glass(int(N, c_int), zJ, jc, jn, jv, sg, sgs, bad)
    if (st /= ESP_OK) then
      write(out_u, '(2x,a)') 'Jij_parameters block is not symmetric.'
      write(out_u, '(2x,a,i4,a,i4,a,5x,a,i4,a,i4,a)') 'Asymmetric pairs are (', bad(1), ',', bad(2), ')', &
                                                     '(', bad(2), ',', bad(1), ')'
      call input_error('Found asymmetric pairs in Jij_parameters block')
    end if
  end subroutine

  ! Optional observables on the device (esp_set_orderparam / _staggered / _pm / _spincorr /
  ! _sfactor) from the input's flags; P(m) is always on, as the reference always writes _pm.dat.
  subroutine setup_observables()
    integer(c_int), allocatable :: kinds(:), scnt(:,:), snbr(:,:,:), c1(:), n1(:)
    real(c_double), allocatable :: f1(:,:), f2(:,:), v(:), val(:), pos(:,:), q(:,:)
    integer(c_int) :: z, zs, m, st, nq
    integer :: sh, i, j, x0, x1, x2, a, lx, ly
    if (inp%lorderparam) call check(esp_set_orderparam(ctx, inp%op_axes), 'esp_set_orderparam')
    if (inp%lstaggered) call check(esp_set_staggered(ctx, 1_c_int, inp%stag_coeff), 'esp_set_staggered')
    call check(esp_set_pm(ctx, 1_c_int, nfreq), 'esp_set_pm')
    if (inp%lspincorr) then
      ! shell tables: the J rows of each shell expanded alone (shell_neighbors_get,
      ! jij.F90:449-503, reads jij_get's slots labelled by the rows' sh=)
      nshell = maxval(inp%J%shell)
      zs = 1
      allocate(scnt(N, nshell), snbr(N, inp%J%n, nshell))
      scnt = 0
      snbr = 0
      do sh = 1, nshell
        m = count(inp%J%shell == sh)
        if (m == 0) cycle
        allocate(f1(3, m), f2(3, m), v(m), c1(N), n1(N * m), val(N * m))
        j = 0
        do i = 1, inp%J%n
          if (inp%J%shell(i) /= sh) cycle
          j = j + 1
          f1(:, j) = inp%J%f1(:, i)
          f2(:, j) = inp%J%f2(:, i)
        end do
        v = 1
        st = esp_build_table(int(inp%nb, c_int), inp%basis, inp%moment, inp%L, &
                             merge(1_c_int, 0_c_int, inp%open_boundary), m, f1, f2, v, c_null_ptr, 1_c_int, &
                             m, c1, n1, val, z)
        if (st /= ESP_OK) call input_error('Unrecognised Atomic Position found in jij_parameters')
        scnt(:, sh) = c1
        do j = 1, z
          snbr(:, j, sh) = n1((j - 1) * N + 1:j * N)
        end do
        zs = max(zs, z)
        deallocate(f1, f2, v, c1, n1, val)
      end do
      allocate(kinds(N))
      do i = 1, N
        kinds(i) = inp%kind(mod(i - 1, inp%nb) + 1) - 1
      end do
      call check(esp_set_spincorr(ctx, int(nshell, c_int), int(inp%nsp, c_int), kinds, zs, scnt, &
                                  pack_nbr(snbr, zs)), 'esp_set_spincorr')
    end if
    if (inp%lsfactor) then
      ! q-points of the 2-D mesh (mcarlo_structure_factor, get_quant.F90:671-681) and the
      ! Cartesian site positions (mcarlo_get_atoms_pos)
      nq = (inp%sf_mesh(1) + 1) * (inp%sf_mesh(2) + 1)
      allocate(q(3, nq), pos(3, N))
      do i = 0, nq - 1
        lx = i / (inp%sf_mesh(2) + 1)
        ly = i - lx * (inp%sf_mesh(2) + 1)
        q(:, i + 1) = inp%sf_corner + lx / real(inp%sf_mesh(1), dp) * inp%sf_q1 + &
                      ly / real(inp%sf_mesh(2), dp) * inp%sf_q2
      end do
      i = 0
      do x2 = 0, inp%L(3) - 1
        do x1 = 0, inp%L(2) - 1
          do x0 = 0, inp%L(1) - 1
            do a = 1, inp%nb
              i = i + 1
              pos(:, i) = matmul(inp%lat, [real(x0, dp), real(x1, dp), real(x2, dp)] + inp%basis(:, a))
            end do
          end do
        end do
      end do
      call check(esp_set_sfactor(ctx, nq, q, pos, inp%sf_polar, int(inp%sf_steps_measure, c_int)), &
                 'esp_set_sfactor')
    end if
  end subroutine

  ! nbr(N, z, shell) with the realised width zs (the C layout [shell][zs][N])
  function pack_nbr(nb3, zs) result(out)
    integer(c_int), intent(in) :: nb3(:,:,:)
    integer(c_int), intent(in) :: zs
    integer(c_int), allocatable :: out(:)
    integer :: sh, l
    allocate(out(size(nb3, 1) * zs * size(nb3, 3)))
    do sh = 1, size(nb3, 3)
      do l = 1, zs
        out(((sh - 1) * zs + (l - 1)) * size(nb3, 1) + 1:((sh - 1) * zs + l) * size(nb3, 1)) = nb3(:, l, sh)
      end do
    end do
  end function

  ! _pm.dat, _op.dat, _staggered.dat, _spincorr.dat, _binerror.dat, _sf-TNNN.dat of the
  ! measured phase (mcarlo.F90:713-820, pt.F90:1000-1110), values by temperature slot.
  subroutine write_observables(pt, nm)
    logical, intent(in) :: pt
    integer, intent(in) :: nm
    real(c_double), allocatable :: pm(:,:,:), ob(:,:), sc(:,:,:,:,:), sf(:,:,:), tr(:,:,:), otr(:,:,:)
    real(dp), allocatable :: ser(:,:), var(:,:)
    integer(c_int64_t) :: cnt
    integer(c_int) :: nrec
    integer :: u, t, k, nd, ns, l, a, b, nq, nk
    real(dp) :: nrm, mmax, dmm
    nrm = 1.0_dp / real(nm, dp)
    ! ---- P(m) (Pm_get, get_quant.F90:425-493; normalisation mcarlo.F90:715) ----------------
    nd = 1 + merge(1, 0, inp%lorderparam) + merge(1, 0, inp%lstaggered)
    allocate(pm(nfreq, nd, R))
    call check(esp_get_obs(ctx, ESP_OBS_PM, pm, cnt), 'esp_get_obs')
    mmax = sum(abs(inp%moment))
    dmm = 2 * mmax / real(nfreq - 1, dp)
    open(newunit=u, file=trim(seedname)//'_pm.dat', status='replace', action='write')
    write(u, *) '#written on '//stamp()
    call pm_header(u, pt)
    do t = 1, R
      write(u, '(1x,a,2x,f9.4)') '#T=', tems(t)
      do k = 1, nfreq
        write(u, '(2x,E15.8,4(10x,E15.8))') -mmax + real(k - 1, dp) * dmm, &
          (pm(k, a, t) * (1.0_dp / 0.01_dp) * nrm / real(N / inp%nb, dp), a = 1, nd)
      end do
    end do
    close(u)
    ! ---- order parameter / staggered moment (mcarlo.F90:773-784) ---------------------------
    allocate(ob(3, R))
    if (inp%lorderparam) then
      call check(esp_get_obs(ctx, ESP_OBS_ORDER, ob, cnt), 'esp_get_obs')
      call moment_file('_op.dat', '#------------------------------- MONTECARLO -----------------------------------#', &
                       '#    Temp              OP                 Sus_OP                 U_OP          #', &
                       '#------------------------------------------------------------------------------#', ob, nrm, pt)
    end if
    if (inp%lstaggered) then
      call check(esp_get_obs(ctx, ESP_OBS_STAGGERED, ob, cnt), 'esp_get_obs')
      call moment_file('_staggered.dat', '#------------------------------- MONTECARLO --------------------------------#', &
                       '#    Temp         Staggered_m          Sus_Staggered_m       U_Staggered_m  #', &
                       '#---------------------------------------------------------------------------#', ob, nrm, pt)
    end if
    ! ---- spin correlations (mcarlo.F90:753-764) ---------------------------------------------
    if (inp%lspincorr) then
      ns = inp%nsp
      allocate(sc(ns, ns, nshell, 3, R))
      call check(esp_get_obs(ctx, ESP_OBS_SPINCORR, sc, cnt), 'esp_get_obs')
      open(newunit=u, file=trim(seedname)//'_spincorr.dat', status='replace', action='write')
      write(u, *) '#written on '//stamp()
      write(u, '(1x,a)') '#---------------------------------------------------- MONTECARLO '// &
                         '----------------------------------------------------#'
      write(u, '(1x,a)') '#  Temp    Shell   Atom1   Type1   Atom2   Type2   '// &
                         '<\sum_{ij} Si.Sj/N>   <\sum_{ij}|Si.Sj|/N>   <|\sum_{ij}Si.Sj|/N> #'
      write(u, '(1x,a)') '#----------------------------------------------------------------'// &
                         '----------------------------------------------------#'
      do t = 1, R
        do l = 1, nshell
          do a = 1, ns
            do b = 1, ns
              write(u, '(f9.4,1x,I5,6x,a3,4x,i3,6x,a3,4x,i3,8x,f12.9,10x,f12.9,12x,f12.9)') tems(t), l, &
                symbol(a), a, symbol(b), b, sc(b, a, l, 1, t) * nrm / N, sc(b, a, l, 3, t) * nrm / N, &
                sc(b, a, l, 2, t) * nrm / N
            end do
          end do
        end do
      end do
      close(u)
    end if
    ! ---- binning errors (mcarlo.F90:548-590,635-656; pt.F90:840-932) ------------------------
    if (inp%lbinerror) then
      allocate(tr(2, R, nm))
      call check(esp_get_trace(ctx, tr, nrec), 'esp_get_trace')
      nk = 2 + merge(1, 0, inp%lorderparam) + merge(1, 0, inp%lstaggered)
      if (nk > 2) then
        allocate(otr(2, R, nm))
        call check(esp_get_obs_trace(ctx, otr, nrec), 'esp_get_obs_trace')
      end if
      open(newunit=u, file=trim(seedname)//'_binerror.dat', status='replace', action='write')
      write(u, *) '#written on '//stamp()
      call binerror_header(u, pt)
      allocate(ser(nm, nk), var(nk, size(inp%binning_level)))
      do t = 1, R
        ser(:, 1) = tr(1, t, 1:nm) / real(N, dp)
        ser(:, 2) = tr(2, t, 1:nm)
        k = 2
        if (inp%lorderparam) then
          k = k + 1
          ser(:, k) = otr(1, t, 1:nm)
        end if
        if (inp%lstaggered) then
          k = k + 1
          ser(:, k) = otr(2, t, 1:nm)
        end if
        call binning(ser, inp%binning_level, pt, var)
        if (pt) then
          write(u, '(1x,a,2x,f6.2)') '#T=', tems(t)
        else
          write(u, '(1x,a,2x,f9.4)') '#T=', tems(t)
        end if
        do l = 1, size(inp%binning_level)
          if (nk == 4 .and. .not. pt) then
            write(u, '(2x,I3,8x,E15.8,3(3x,E15.8))') inp%binning_level(l), (sqrt(var(a, l) / N), a = 1, nk)
          else
            write(u, '(2x,I8,8x,E15.8,3(3x,E15.8))') inp%binning_level(l), (sqrt(var(a, l) / N), a = 1, nk)
          end if
        end do
      end do
      close(u)
    end if
    ! ---- structure factor (get_quant.F90:719-920) -----------------------------------------
    if (inp%lsfactor) then
      nq = (inp%sf_mesh(1) + 1) * (inp%sf_mesh(2) + 1)
      allocate(sf(nq, 3, R))
      call check(esp_get_obs(ctx, ESP_OBS_SFACTOR, sf, cnt), 'esp_get_obs')
      call write_sfactor(sf, cnt)
    end if
  end subroutine

  character(len=24) function stamp()
    character(len=8) :: d
    character(len=10) :: tm
    call date_and_time(date=d, time=tm)
    stamp = d(7:8)//'/'//d(5:6)//'/'//d(1:4)//' at '//tm(1:2)//':'//tm(3:4)//':'//tm(5:6)
  end function

  ! the reference's atoms_symbol: the species label with an upper-case first letter
  character(len=3) function symbol(k)
    integer, intent(in) :: k
    integer :: c
    symbol = inp%species(k)
    c = ichar(symbol(1:1))
    if (c >= ichar('a') .and. c <= ichar('z')) symbol(1:1) = char(c - 32)
  end function

  ! _op.dat / _staggered.dat: <|X|>, susceptibility, Binder cumulant per temperature
  ! (mcarlo.F90:773-784, format 126 of mcarlo.F90:855 / pt.F90:1128)
  subroutine moment_file(suffix, h1, h2, h3, ob, nrm, pt)
    character(len=*), intent(in) :: suffix, h1, h2, h3
    real(c_double), intent(in) :: ob(:,:)
    real(dp), intent(in) :: nrm
    logical, intent(in) :: pt
    integer :: uu, tt
    real(dp) :: bt, x1, x2, x4
    open(newunit=uu, file=trim(seedname)//suffix, status='replace', action='write')
    write(uu, *) '#written on '//stamp()
    write(uu, '(1x,a)') h1
    write(uu, '(1x,a)') h2
    write(uu, '(1x,a)') h3
    do tt = 1, R
      bt = 1.0_dp / tems(tt)
      x1 = ob(1, tt) * nrm
      x2 = ob(2, tt) * nrm
      x4 = ob(3, tt) * nrm
      if (pt) then
        write(uu, '(3x,f9.4,5x,E15.8,6x,E15.8,8x,E15.8,5x)') tems(tt), x1, bt * (x2 - x1**2) * N, &
                                                             1.0_dp - (1.0_dp / 3.0_dp) * (x4 / x2**2)
      else
        write(uu, '(3x,f9.4,5x,E15.8,6x,E15.8,7x,E15.8,5x)') tems(tt), x1, bt * (x2 - x1**2) * N, &
                                                             1.0_dp - (1.0_dp / 3.0_dp) * (x4 / x2**2)
      end if
    end do
    close(uu)
  end subroutine

  ! table headers of _pm.dat and _binerror.dat (mcarlo.F90:172-215, pt.F90:159-200)
  subroutine pm_header(u, pt)
    integer, intent(in) :: u
    logical, intent(in) :: pt
    if (inp%lorderparam .and. inp%lstaggered) then
      if (pt) then
        write(u, '(1x,a)') '#-------------------------------------- MONTECARLO ------'//'------------------------------------#'
        write(u, '(1x,a)') '#       M or OP                 P(M)                     '//'P(OP)               P(Staggered_m)  #'
        write(u, '(1x,a)') '#--------------------------------------------------------'//'------------------------------------#'
      else
        write(u, '(1x,a)') '#------------------------------------- MONTECARLO ---------'//'----------------------------------#'
        write(u, '(1x,a)') '#       M or OP                 P(M)                     P('//'OP)               P(Staggered_m)  #'
        write(u, '(1x,a)') '#----------------------------------------------------------'//'----------------------------------#'
      end if
    else if (inp%lorderparam) then
      if (pt) then
        write(u, '(1x,a)') '#---------------------------- MONTECARLO -----------------------------#'
      else
        write(u, '(1x,a)') '#-------------------------- MONTECARLO -------------------------------#'
      end if
      write(u, '(1x,a)') '#       M or OP                 P(M)                     P(OP)        #'
      write(u, '(1x,a)') '#---------------------------------------------------------------------#'
    else if (inp%lstaggered) then
      if (pt) then
        write(u, '(1x,a)') '#--------------------------- MONTECARLO-----------------------------#'
      else
        write(u, '(1x,a)') '#-------------------------- MONTECARLO -----------------------------#'
      end if
      write(u, '(1x,a)') '# M or Staggered                P(M)                P(Staggered_m)  #'
      write(u, '(1x,a)') '#-------------------------------------------------------------------#'
    else
      write(u, '(1x,a)') '#----------------- MONTECARLO -------------------#'
      write(u, '(1x,a)') '#        M                      P(M)             #'
      write(u, '(1x,a)') '#------------------------------------------------#'
    end if
  end subroutine

  subroutine binerror_header(u, pt)
    integer, intent(in) :: u
    logical, intent(in) :: pt
    if (inp%lorderparam .and. inp%lstaggered) then
      if (pt) then
        write(u, '(1x,a)') '#------------------------------------- MONTECARLO ---------------------'//'--------------------#'
        write(u, '(1x,a)') '#  Binning_level     Error(E)          Error(M)        Error(OP)       '//'Error(Staggered_m)  #'
        write(u, '(1x,a)') '#----------------------------------------------------------------------'//'--------------------#'
      else
        write(u, '(1x,a)') '#------------------------------------ MONTECARLO ------'//'------------------------------------#'
        write(u, '(1x,a)') '#  Binning_level     Error(E)          Error(M)        '//'Error(OP)       Error(Staggered_m)  #'
        write(u, '(1x,a)') '#------------------------------------------------------'//' ------------------------------------#'
      end if
    else if (inp%lorderparam) then
      if (pt) then
        write(u, '(1x,a)') '#--------------------------- MONTECARLO -----------------------------#'
      else
        write(u, '(1x,a)') '#------------------------- MONTECARLO -------------------------------#'
      end if
      write(u, '(1x,a)') '#  Binning_level     Error(E)          Error(M)         Error(OP)    #'
      write(u, '(1x,a)') '#--------------------------------------------------------------------#'
    else if (inp%lstaggered) then
      write(u, '(1x,a)') '#----------------------------- MONTECARLO -------------------------------#'
      write(u, '(1x,a)') '#  Binning_level     Error(E)          Error(M)      Error(Staggered_m)  #'
      write(u, '(1x,a)') '#------------------------------------------------------------------------#'
    else
      write(u, '(1x,a)') '#------------------- MONTECARLO -----------------#'
      write(u, '(1x,a)') '#  Block_length      Error(E)          Error(M)  #'
      write(u, '(1x,a)') '#------------------------------------------------#'
    end if
  end subroutine

  ! Binning analysis of one temperature's measurement series ser(n, k): box means of 2^level
  ! measurements, their spread, the incomplete last box left out (mcarlo.F90:582-590,635-656).
  ! The parallel-tempering loop divides a completed box by 2^(index of the level) instead of
  ! 2^level (pt.F90:843); with the default levels 1..5 both agree.
  subroutine binning(ser, levels, pt, var)
    real(dp), intent(in) :: ser(:,:)
    integer, intent(in) :: levels(:)
    logical, intent(in) :: pt
    real(dp), intent(out) :: var(:,:)
    real(dp) :: box(size(ser, 2), size(levels)), acc2(size(ser, 2), size(levels)), tot(size(ser, 2))
    integer :: n, st, l, ml
    n = size(ser, 1)
    box = 0
    acc2 = 0
    tot = 0
    do st = 1, n
      tot = tot + ser(st, :)
      do l = 1, size(levels)
        box(:, l) = box(:, l) + ser(st, :)
        if (mod(st, 2**levels(l)) == 0) then
          if (pt) then
            box(:, l) = box(:, l) / real(2**l, dp)
          else
            box(:, l) = box(:, l) / real(2**levels(l), dp)
          end if
          acc2(:, l) = acc2(:, l) + box(:, l)**2
          box(:, l) = 0
        end if
      end do
    end do
    do l = 1, size(levels)
      ml = n / 2**levels(l)
      box(:, l) = (tot - box(:, l)) / real(2**levels(l), dp) / real(ml, dp)
      var(:, l) = (acc2(:, l) / real(ml, dp) - box(:, l)**2) / real(ml - 1, dp)
    end do
  end subroutine

  ! seedname_sf-TNNN.dat per temperature: in-plane coordinates of every q-point of the mesh
  ! and S(q) per site and measurement (mcarlo_structure_factor_write_mc, get_quant.F90:719-820),
  ! plus a short plotting script.
  subroutine write_sfactor(sf, cnt)
    real(c_double), intent(in) :: sf(:,:,:)
    integer(c_int64_t), intent(in) :: cnt
    real(dp), parameter :: twopi = 6.283185307179586476925286766559_dp
    real(dp) :: qv(3,3), av(3,3), zv(3), yv(3), q1m, q2m, ym, cq, cy, c1, c2, qx, qy, f1, f2, e(3)
    integer :: u, t, i, lx, ly, k
    character(len=300) :: fn
    q1m = sqrt(sum(inp%sf_q1**2))
    q2m = sqrt(sum(inp%sf_q2**2))
    zv = cross(inp%sf_q1, inp%sf_q2)
    yv = cross(zv, inp%sf_q1)
    ym = sqrt(sum(yv**2))
    qv(:, 1) = inp%sf_q1
    qv(:, 2) = inp%sf_q2
    qv(:, 3) = zv / sqrt(sum(zv**2))
    ! dual vectors: av(:, k) . qv(:, j) = 2 pi delta_jk
    do k = 1, 3
      e = 0
      e(k) = 1
      block
        real(dp) :: f(3)
        call solve3(qv, e, f)
        av(k, :) = twopi * f
      end block
    end do
    cy = dot_product(yv, inp%sf_q2) / (ym * q2m)
    cq = dot_product(inp%sf_q1, inp%sf_q2) / (q1m * q2m)
    c1 = dot_product(inp%sf_corner, av(:, 1)) / twopi
    c2 = dot_product(inp%sf_corner, av(:, 2)) / twopi
    do t = 1, R
      write(fn, '(a,a,i3.3,a)') trim(seedname), '_sf-T', t, '.dat'
      write(out_u, '(3x,a,f8.3,3x,a,3x,a)') 'Output Structure factor at T = ', tems(t), 'writted at', trim(fn)
      open(newunit=u, file=trim(fn), status='replace', action='write')
      do i = 0, size(sf, 1) - 1
        lx = i / (inp%sf_mesh(2) + 1)
        ly = i - lx * (inp%sf_mesh(2) + 1)
        f1 = lx / real(inp%sf_mesh(1), dp) + c1
        f2 = ly / real(inp%sf_mesh(2), dp) + c2
        qx = f1 * q1m + f2 * q2m * cq
        qy = f2 * q2m * cy
        if (inp%lsf_polar) then
          write(u, '(3x,E15.8,3x,E15.8,3x,E15.8,3x,E15.8,3x,E15.8)') qx, qy, &
            (sf(i + 1, k, t) / N / real(max(cnt, 1_c_int64_t), dp), k = 1, 3)
        else
          write(u, '(3x,E15.8,3x,E15.8,3x,E15.8)') qx, qy, sf(i + 1, 1, t) / N / real(max(cnt, 1_c_int64_t), dp)
        end if
      end do
      close(u)
    end do
    write(out_u, '(4x,a,I14,17x,a)') ' Total used steps at S Factor calculation = ', cnt, '   <-- SFAC'
    open(newunit=u, file=trim(seedname)//'_sf.py', status='replace', action='write')
    write(u, '(a)') '# S(q) of the first temperature on the (q1, q2) plane: columns qx, qy, S'
    write(u, '(a)') 'import numpy as np, matplotlib.pyplot as plt'
    write(u, '(a)') "d = np.loadtxt('"//trim(seedname)//"_sf-T001.dat')"
    write(u, '(a)') 'plt.tricontourf(d[:, 0], d[:, 1], d[:, 2], 200)'
    write(u, '(a)') "plt.gca().set_aspect('equal'); plt.colorbar()"
    write(u, '(a)') "plt.savefig('"//trim(seedname)//"_sf-T001.pdf')"
    close(u)
  end subroutine

  function cross(a, b) result(c)
    real(dp), intent(in) :: a(3), b(3)
    real(dp) :: c(3)
    c = [a(2) * b(3) - a(3) * b(2), a(3) * b(1) - a(1) * b(3), a(1) * b(2) - a(2) * b(1)]
  end function

  ! x with A^T x = e  (A's columns are the vectors; x is then a dual vector over 2 pi)
  subroutine solve3(A, e, x)
    real(dp), intent(in) :: A(3,3), e(3)
    real(dp), intent(out) :: x(3)
    real(dp) :: M(3,3), det
    M = transpose(A)
    det = M(1,1)*(M(2,2)*M(3,3)-M(3,2)*M(2,3)) - M(1,2)*(M(2,1)*M(3,3)-M(2,3)*M(3,1)) &
        + M(1,3)*(M(2,1)*M(3,2)-M(2,2)*M(3,1))
    x(1) = (e(1)*(M(2,2)*M(3,3)-M(3,2)*M(2,3)) - M(1,2)*(e(2)*M(3,3)-M(2,3)*e(3)) &
          + M(1,3)*(e(2)*M(3,2)-M(2,2)*e(3))) / det
    x(2) = (M(1,1)*(e(2)*M(3,3)-e(3)*M(2,3)) - e(1)*(M(2,1)*M(3,3)-M(2,3)*M(3,1)) &
          + M(1,3)*(M(2,1)*e(3)-e(2)*M(3,1))) / det
    x(3) = (M(1,1)*(M(2,2)*e(3)-M(3,2)*e(2)) - M(1,2)*(M(2,1)*e(3)-e(2)*M(3,1)) &
          + e(1)*(M(2,1)*M(3,2)-M(2,2)*M(3,1))) / det
  end subroutine

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
