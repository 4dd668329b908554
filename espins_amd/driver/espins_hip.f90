! espins_hip.f90 -- ISO_C_BINDING interface to libespins_hip.so (include/espins_hip.h).
!
! Every entry point returns an integer(c_int) status: ESP_OK = 0, negative on error, with the
! message from esp_last_error.  Arrays are passed as the reference's own Fortran arrays:
! spin_matrix(3,N,R), jij_matrix(N,z), jij_nbors_matrix(N,z) (1-based), dij_matrix(3,N,z).
module espins_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ESP_OK = 0, ESP_MODE_RANDOM = 0, ESP_MODE_CONST = 1, &
                                       ESP_MODE_LEARNED = 2, ESP_LEARN_NPARAMS = 2502

  public :: esp_create, esp_destroy, esp_last_error, esp_version, esp_build_table, &
            esp_set_supercell, esp_set_model, esp_get_model_info, esp_set_replicas, &
            esp_get_spins, esp_put_spins, esp_energy, esp_run, esp_get_stats, esp_get_map, &
            esp_reset_stats, esp_sync, esp_error_text, esp_sweep, esp_learn_set_weights, &
            esp_learn_collect, esp_learn_train, esp_learn_clear, esp_set_trace, esp_get_trace, &
            esp_set_orderparam, esp_set_staggered, esp_set_pm, esp_set_spincorr, esp_set_sfactor, &
            esp_get_obs, esp_get_obs_trace, esp_spin_glass

  integer(c_int), parameter, public :: ESP_OBS_ORDER = 1, ESP_OBS_STAGGERED = 2, ESP_OBS_PM = 4, &
                                       ESP_OBS_SPINCORR = 8, ESP_OBS_SFACTOR = 16

  interface
    integer(c_int) function esp_create(ctx, device) bind(C, name='esp_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function
    integer(c_int) function esp_destroy(ctx) bind(C, name='esp_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function esp_last_error(ctx) bind(C, name='esp_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function esp_version() bind(C, name='esp_version')
      import :: c_int
    end function
    integer(c_int) function esp_build_table(nb, basis_frac, basis_mag, L, open_boundary, nrows, &
        f1, f2, value, dvec, ncomp, zcap, cnt, nbr, val, zmax) bind(C, name='esp_build_table')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nb, open_boundary, nrows, ncomp, zcap
      real(c_double), intent(in) :: basis_frac(*), basis_mag(*), f1(*), f2(*), value(*)
      type(c_ptr), value :: dvec
      integer(c_int), intent(in) :: L(3)
      integer(c_int), intent(out) :: cnt(*), nbr(*), zmax
      real(c_double), intent(out) :: val(*)
    end function
    integer(c_int) function esp_set_supercell(ctx, L) bind(C, name='esp_set_supercell')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: L(3)
    end function
    integer(c_int) function esp_set_model(ctx, N, nb, mag, zJ, j_cnt, j_nbr, j_val, &
        zB, b_cnt, b_nbr, b_val, zD, d_cnt, d_nbr, d_vec, have_si, si_par, si_axis, &
        have_field, h_par, h_axis) bind(C, name='esp_set_model')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: N, nb, zJ, zB, zD, have_si, have_field
      real(c_double), intent(in) :: mag(*)
      integer(c_int), intent(in) :: j_cnt(*), j_nbr(*)
      real(c_double), intent(in) :: j_val(*)
      type(c_ptr), value :: b_cnt, b_nbr, b_val, d_cnt, d_nbr, d_vec, si_par, si_axis, h_par, h_axis
    end function
    integer(c_int) function esp_get_model_info(ctx, out6) bind(C, name='esp_get_model_info')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: out6(6)
    end function
    integer(c_int) function esp_set_replicas(ctx, R, tems, spins) bind(C, name='esp_set_replicas')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: R
      real(c_double), intent(in) :: tems(*), spins(*)
    end function
    integer(c_int) function esp_get_spins(ctx, by_slot, spins) bind(C, name='esp_get_spins')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: by_slot
      real(c_double), intent(out) :: spins(*)
    end function
    integer(c_int) function esp_put_spins(ctx, spins) bind(C, name='esp_put_spins')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: spins(*)
    end function
    integer(c_int) function esp_energy(ctx, E) bind(C, name='esp_energy')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: E(*)
    end function
    integer(c_int) function esp_run(ctx, n_rounds, steps_per_round, steps_measure, do_measure, &
        do_exchange, start_parity, mode, tilt, seed, sweep0, step0, use_graph) bind(C, name='esp_run')
      import :: c_int, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_rounds, steps_per_round, steps_measure, do_measure, do_exchange, &
                               start_parity, mode, use_graph
      real(c_double), value :: tilt
      integer(c_int64_t), value :: seed, sweep0, step0
    end function
    integer(c_int) function esp_get_stats(ctx, acc, rejected, e_run, nmeas) bind(C, name='esp_get_stats')
      import :: c_int, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: acc(*), e_run(*)
      integer(c_int64_t), intent(out) :: rejected(*), nmeas
    end function
    integer(c_int) function esp_get_map(ctx, slot_rep, nswap) bind(C, name='esp_get_map')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: slot_rep(*)
      integer(c_int64_t), intent(out) :: nswap(*)
    end function
    integer(c_int) function esp_reset_stats(ctx) bind(C, name='esp_reset_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function esp_sync(ctx) bind(C, name='esp_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function esp_sweep(ctx, nsweeps, mode, tilt, seed, sweep0) bind(C, name='esp_sweep')
      import :: c_int, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsweeps, mode
      real(c_double), value :: tilt
      integer(c_int64_t), value :: seed, sweep0
    end function
    integer(c_int) function esp_set_trace(ctx, capacity) bind(C, name='esp_set_trace')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: capacity
    end function
    integer(c_int) function esp_get_trace(ctx, trace, n) bind(C, name='esp_get_trace')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: trace(*)
      integer(c_int), intent(out) :: n
    end function
    integer(c_int) function esp_spin_glass(N, z, cnt, nbr, val, sigma, seed, bad) bind(C, name='esp_spin_glass')
      import :: c_int, c_double
      integer(c_int), value :: N, z, seed
      integer(c_int), intent(in) :: cnt(*), nbr(*)
      real(c_double), intent(inout) :: val(*)
      real(c_double), intent(in) :: sigma(*)
      integer(c_int), intent(out) :: bad(2)
    end function
    ! optional observables
    integer(c_int) function esp_set_orderparam(ctx, axes) bind(C, name='esp_set_orderparam')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: axes(*)
    end function
    integer(c_int) function esp_set_staggered(ctx, enable, coeff) bind(C, name='esp_set_staggered')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      real(c_double), intent(in) :: coeff(*)
    end function
    integer(c_int) function esp_set_pm(ctx, enable, nfreq) bind(C, name='esp_set_pm')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      integer(c_int), intent(out) :: nfreq
    end function
    integer(c_int) function esp_set_spincorr(ctx, nshell, nspecies, species, zs, cnt, nbr) &
        bind(C, name='esp_set_spincorr')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nshell, nspecies, zs
      integer(c_int), intent(in) :: species(*), cnt(*), nbr(*)
    end function
    integer(c_int) function esp_set_sfactor(ctx, nq, q, pos, polar, every) bind(C, name='esp_set_sfactor')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nq, every
      real(c_double), intent(in) :: q(*), pos(*), polar(*)
    end function
    integer(c_int) function esp_get_obs(ctx, which, out, count) bind(C, name='esp_get_obs')
      import :: c_int, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function esp_get_obs_trace(ctx, trace, n) bind(C, name='esp_get_obs_trace')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: trace(*)
      integer(c_int), intent(out) :: n
    end function
    ! learned proposal (no counterpart in the reference)
    integer(c_int) function esp_learn_set_weights(ctx, w, nparams) bind(C, name='esp_learn_set_weights')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: w(*)
      integer(c_int), value :: nparams
    end function
    integer(c_int) function esp_learn_collect(ctx, n, seed) bind(C, name='esp_learn_collect')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function esp_learn_clear(ctx) bind(C, name='esp_learn_clear')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function esp_learn_train(ctx, epochs, batch, lr, seed, epoch_loss) &
        bind(C, name='esp_learn_train')
      import :: c_int, c_float, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: epochs, batch
      real(c_float), value :: lr
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: epoch_loss(*)
    end function
  end interface

contains

  ! The library's message for the last failed call on ctx.
  function esp_error_text(ctx) result(txt)
    type(c_ptr), value :: ctx
    character(len=512) :: txt
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: k
    txt = ' '
    p = esp_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do k = 1, 512
      if (s(k) == c_null_char) exit
      txt(k:k) = s(k)
    end do
  end function

end module espins_hip
