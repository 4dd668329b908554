! mc_echo.f90 -- the input echo at the top of seedname.mcout and the random-number box before
! the Monte Carlo tables, in the layout of param_write (parameters.F90:1074-1444) and
! param_write_seed (parameters.F90:1447-1494): the same sections, column widths, unit labels
! and conversions (length_unit scales lengths by 1/bohr, parameter_unit = ryd prints couplings
! times 2/hart), so a reader of the reference's .mcout reads this one.
module mc_echo
  use mc_input
  implicit none
  private
  public :: write_echo, write_seed_box

  real(dp), parameter :: twopi = 6.283185307179586476925286766559_dp
  integer, parameter :: W = 85
  character(len=W), parameter :: rule = '*' // repeat('-', W - 2) // '*'
  character(len=W), parameter :: plus = '+' // repeat('-', W - 2) // '+'

contains

  ! "*---- title ----*", left dashes as the reference places each title
  subroutine title(u, text, left)
    integer, intent(in) :: u, left
    character(len=*), intent(in) :: text
    write(u, '(1x,a)') '*' // repeat('-', left) // ' ' // text // ' ' // repeat('-', W - 4 - left - len(text)) // '*'
  end subroutine

  ! "|<text>      |" over the full width (a text longer than the box is cut at 85 columns, as
  ! an a85 edit descriptor cuts it)
  subroutine boxed(u, text)
    integer, intent(in) :: u
    character(len=*), intent(in) :: text
    character(len=W) :: line
    line = '|' // text
    if (len(text) < W - 1) line(W:W) = '|'
    write(u, '(1x,a)') line
  end subroutine

  ! key/value rows: "|  <label>  :" in 48 columns, the value right-aligned in 9 after 10 blanks
  function key48(label) result(k)
    character(len=*), intent(in) :: label
    character(len=48) :: k
    character(len=43) :: l
    l = label
    k = '|  ' // l // ':'
  end function

  subroutine kv_l(u, label, v)
    integer, intent(in) :: u
    character(len=*), intent(in) :: label
    logical, intent(in) :: v
    write(u, '(1x,a48,10x,l9,17x,a1)') key48(label), v, '|'
  end subroutine

  subroutine kv_i(u, label, v)
    integer, intent(in) :: u, v
    character(len=*), intent(in) :: label
    write(u, '(1x,a48,10x,i9,17x,a1)') key48(label), v, '|'
  end subroutine

  subroutine kv_f(u, label, v)
    integer, intent(in) :: u
    character(len=*), intent(in) :: label
    real(dp), intent(in) :: v
    write(u, '(1x,a48,10x,f9.3,17x,a1)') key48(label), v, '|'
  end subroutine

  subroutine kv_a(u, label, v)
    integer, intent(in) :: u
    character(len=*), intent(in) :: label, v
    write(u, '(1x,a48,10x,a9,17x,a1)') key48(label), trim(v), '|'
  end subroutine

  ! exact zeros printed without a sign
  elemental function nz(x) result(y)
    real(dp), intent(in) :: x
    real(dp) :: y
    y = merge(0.0_dp, x, x == 0.0_dp)
  end function

  function capital(s) result(t)
    character(len=*), intent(in) :: s
    character(len=3) :: t
    integer :: c
    t = s
    c = iachar(t(1:1))
    if (c >= iachar('a') .and. c <= iachar('z')) t(1:1) = achar(c - 32)
  end function

  ! fractional coordinates of a Cartesian real-space vector (rows of recip / 2 pi)
  function frac_of(inp, v) result(f)
    type(input_t), intent(in) :: inp
    real(dp), intent(in) :: v(3)
    real(dp) :: f(3)
    integer :: i
    do i = 1, 3
      f(i) = dot_product(inp%recip(:, i), v) / twopi
    end do
  end function

  ! reduced coordinates of a Cartesian reciprocal-space vector (lattice vectors / 2 pi)
  function reduced_of(inp, q) result(f)
    type(input_t), intent(in) :: inp
    real(dp), intent(in) :: q(3)
    real(dp) :: f(3)
    integer :: i
    do i = 1, 3
      f(i) = dot_product(inp%lat(:, i), q) / twopi
    end do
  end function

  subroutine coordinate_line(u, inp)
    integer, intent(in) :: u
    type(input_t), intent(in) :: inp
    if (.not. inp%coord_cart) then
      call boxed(u, '   In Fractional Coordinate:')
    else if (inp%lenconfac == 1.0_dp) then
      call boxed(u, '   In Cartesian Coordinate(Ang):' // repeat(' ', 55) // '|')
    else
      call boxed(u, '   In Cartesian Coordinate(Bohr):' // repeat(' ', 55) // '|')
    end if
  end subroutine

  ! one coupling table (Exchange, Sigma, Biquadratic, Dzyaloshinskii-Moriya, Shell): atom 1 and
  ! atom 2 of every row, then the value (or the shell)
  subroutine pair_table(u, inp, r, name, left, head, vals, shells)
    integer, intent(in) :: u, left
    type(input_t), intent(in) :: inp
    type(rows_t), intent(in) :: r
    character(len=*), intent(in) :: name, head
    real(dp), intent(in), optional :: vals(:)
    integer, intent(in), optional :: shells(:)
    real(dp) :: a(3), b(3)
    integer :: k
    call title(u, name, left)
    call coordinate_line(u, inp)
    write(u, '(1x,a)') rule
    call boxed(u, '              ATOM 1                             ATOM 2' // head)
    write(u, '(1x,a)') rule
    do k = 1, r%n
      if (inp%coord_cart) then
        a = matmul(inp%lat, r%f1(:, k)) * inp%lenconfac
        b = matmul(inp%lat, r%f2(:, k))      ! the reference scales atom 1 only
      else
        a = r%f1(:, k)
        b = r%f2(:, k)
      end if
      if (present(vals)) then
        write(u, '(1x,a1,3f11.6,1x,a1,3f11.6,1x,a,1f12.8,1x,a1)') '|', nz(a), '|', nz(b), '|', vals(k), '|'
      else
        write(u, '(1x,a1,3f11.6,1x,a1,3f11.6,1x,a,1i7,6x,a1)') '|', nz(a), '|', nz(b), '|', shells(k), '|'
      end if
    end do
    write(u, '(1x,a)') rule
  end subroutine

  subroutine write_echo(u, inp)
    integer, intent(in) :: u
    type(input_t), intent(in) :: inp
    real(dp) :: vol, v(3)
    integer :: a, k
    character(len=5) :: eu

    eu = '(eV)'
    if (inp%paramconfac /= 1.0_dp) eu = '(Ryd)'

    ! ---- system --------------------------------------------------------------------------
    write(u, '(a)') ' '
    write(u, '(40x,a6)') '------', 'SYSTEM', '------'
    write(u, '(a)') ' '
    if (inp%lenconfac == 1.0_dp) then
      write(u, '(33x,a)') 'Lattice Vectors (Ang)'
    else
      write(u, '(32x,a)') 'Lattice Vectors (Bohr)'
    end if
    write(u, '(24x,a3,2x,3f11.6)') ('a_' // achar(48 + k), inp%lat(:, k) * inp%lenconfac, k = 1, 3)
    write(u, '(a)') ' '
    vol = abs(inp%lat(1,1) * (inp%lat(2,2) * inp%lat(3,3) - inp%lat(3,2) * inp%lat(2,3)) &
            - inp%lat(1,2) * (inp%lat(2,1) * inp%lat(3,3) - inp%lat(3,1) * inp%lat(2,3)) &
            + inp%lat(1,3) * (inp%lat(2,1) * inp%lat(3,2) - inp%lat(3,1) * inp%lat(2,2)))
    if (inp%lenconfac == 1.0_dp) then
      write(u, '(24x,a17,3x,f11.5,2x,a)') 'Unit Cell Volume:', vol, '(Ang^3)'
    else
      write(u, '(24x,a17,3x,f11.5,2x,a)') 'Unit Cell Volume:', vol * inp%lenconfac**3, '(Bohr^3)'
    end if
    write(u, '(a)') ' '
    if (inp%lenconfac == 1.0_dp) then
      write(u, '(27x,a)') 'Reciprocal-Space Vectors (Ang^-1)'
    else
      write(u, '(27x,a)') 'Reciprocal-Space Vectors (Bohr^-1)'
    end if
    write(u, '(24x,a3,2x,3f11.6)') ('b_' // achar(48 + k), inp%recip(:, k) / inp%lenconfac, k = 1, 3)
    write(u, '(a)') '  '

    ! ---- atoms -----------------------------------------------------------------------------
    write(u, '(1x,a)') rule
    if (inp%lenconfac == 1.0_dp) then
      call boxed(u, ' Atom  Type   Fractional Coordinate             Cartesian Coordinate(Ang)   Moment ')
    else
      call boxed(u, ' Atom  Type   Fractional Coordinate             Cartesian Coordinate(Bohr)  Moment ')
    end if
    write(u, '(1x,a)') plus
    do a = 1, inp%nb
      write(u, '(1x,a1,1x,a3,2x,i3,3f10.5,3x,a1,3f10.5,2x,a1,f6.3,1x,a1)') '|', capital(inp%label(a)), &
        inp%kind(a), nz(inp%basis(:, a)), '|', nz(matmul(inp%lat, inp%basis(:, a)) * inp%lenconfac), '|', inp%moment(a), '|'
    end do
    write(u, '(1x,a)') rule

    ! ---- supercell, Monte Carlo, temperatures, Hamiltonian ----------------------------------
    write(u, '(36x,a)') '--------------', 'SUPERCELL SIZE', '--------------'
    write(u, '(29x,a,i3,1x,a1,i3,1x,a1,i3,6x)') 'Supercell size =', inp%L(1), 'x', inp%L(2), 'x', inp%L(3)
    write(u, '(a)') ' '
    call title(u, 'MONTECARLO', 34)
    call kv_l(u, 'Using Parallel Tempering algorithm', inp%pt)
    call kv_a(u, 'Initial spins configuration', inp%initial)
    call kv_a(u, 'Monte Carlo Mode', inp%mcarlo_mode)
    if (index(inp%mcarlo_mode, 'const') > 0) call kv_f(u, 'Tilting angles maximum', inp%tilt)
    call kv_i(u, 'Number of Warm-up steps', inp%steps_warmup)
    call kv_i(u, 'Number of Monte Carlo sampling steps', inp%steps_mc)
    call kv_i(u, 'Number of steps to take a measurement', inp%steps_measure)
    if (inp%pt) call kv_i(u, 'Number of MC steps to swap configurations', inp%pt_steps_swap)
    call kv_l(u, 'Order parameter (calculate/write)', inp%lorderparam)
    call kv_l(u, 'Staggered Magnetization (calculate/write)', inp%lstaggered)
    call kv_l(u, 'Structure Factor (calculate/write)', inp%lsfactor)
    call kv_l(u, 'Energy               (write)', inp%energy_write)
    call kv_l(u, 'Spin correlation (calculate/write)', inp%lspincorr)
    call kv_l(u, 'Auto correlation (calculate/write)', inp%lbinerror)
    call title(u, 'Temperatures', 33)
    call kv_a(u, 'Temperature mode', inp%tems_mode)
    call kv_i(u, 'Number of temperatures', inp%tems_num)
    if (.not. inp%pt) then
      call kv_f(u, 'Starting temperature', inp%tems(1))
      call kv_f(u, 'Ending temperature', inp%tems(inp%tems_num))
    else
      call kv_f(u, 'Lowest temperature', inp%tems(1))
      call kv_f(u, 'Highest temperature', inp%tems(inp%tems_num))
    end if
    write(u, '(1x,a)') rule
    call title(u, 'Hamiltonian', 34)
    call kv_l(u, 'Exchange', .true.)
    call kv_l(u, 'Using random exchange parameters', inp%spin_glass)
    call kv_l(u, 'Biquadratic', inp%have_B)
    call kv_l(u, 'Dzyaloshinskii-Moriya', inp%have_D)
    call kv_l(u, 'Single-ion', inp%have_si)
    call kv_l(u, 'Magnetic field', inp%have_h)
    write(u, '(1x,a)') rule
    write(u, '(a)') '  '

    ! ---- coupling tables -------------------------------------------------------------------
    if (inp%J%n > 0) then
      call pair_table(u, inp, inp%J, 'Exchange', 36, repeat(' ', 18) // 'J' // trim(eu), vals=inp%J%val * inp%paramconfac)
      write(u, '(a)') '  '
      if (inp%spin_glass) then
        call pair_table(u, inp, inp%J, 'Sigma', 38, repeat(' ', 15) // 'Sigma' // trim(eu), vals=inp%J%sigma * inp%paramconfac)
      else if (inp%lspincorr) then
        call pair_table(u, inp, inp%J, 'Shell', 37, repeat(' ', 19) // 'Shell', shells=inp%J%shell)
      end if
    end if
    if (inp%have_B) then
      call pair_table(u, inp, inp%B, 'Biquadratic', 36, repeat(' ', 18) // 'B' // trim(eu), vals=inp%B%val * inp%paramconfac)
      write(u, '(a)') '  '
    end if
    if (inp%have_D) then
      call pair_table(u, inp, inp%D, 'Dzyaloshinskii-Moriya', 31, repeat(' ', 18) // 'D' // trim(eu), &
                      vals=inp%D%val * inp%paramconfac)
      call boxed(u, ' Moriya-Vector   Fractional Coordinate           Cartesian Coordinate (Normalized) ')
      write(u, '(1x,a)') plus
      do k = 1, inp%D%n
        write(u, '(1x,a1,i6,4x,3f10.5,5x,a1,2x,3f10.5,5x,a1)') '|', k, nz(inp%D%dvec_frac(:, k)), '|', &
          nz(inp%D%dvec(:, k)), '|'
      end do
      write(u, '(1x,a)') rule
      write(u, '(a)') '  '
    end if

    ! ---- axes --------------------------------------------------------------------------------
    if (inp%have_si) then
      call title(u, 'Single-ion-Axes', 34)
      call boxed(u, ' Atoms   Fractional Coordinate     Cartesian Coordinate (Normalized)  \Delta' // trim(eu))
      write(u, '(1x,a)') plus
      do a = 1, inp%nb
        write(u, '(1x,a1,2x,a3,1x,3f8.4,5x,a1,2x,3f8.4,4x,a1,f15.8,1x,a1)') '|', capital(inp%label(a)), &
          nz(frac_of(inp, inp%si_axis(:, a))), '|', nz(inp%si_axis(:, a)), '|', inp%si_par(a) * inp%paramconfac, '|'
      end do
      write(u, '(1x,a)') rule
      write(u, '(a)') '  '
    end if
    if (inp%have_h) then
      call title(u, 'Magnetic Field', 31)
      call boxed(u, ' Atoms   Fractional Coordinate     Cartesian Coordinate (Normalized)      H(T)     ')
      write(u, '(1x,a)') plus
      do a = 1, inp%nb
        write(u, '(1x,a1,2x,a3,1x,3f8.4,2x,a1,1x,3f8.4,7x,a1,1x,f15.8,1x,a1)') '|', capital(inp%label(a)), &
          nz(frac_of(inp, inp%h_axis(:, a))), '|', nz(inp%h_axis(:, a)), '|', inp%h_par(a), '|'
      end do
      write(u, '(1x,a)') rule
      write(u, '(a)') '  '
    end if
    if (inp%lorderparam) then
      call title(u, 'Order-Parameters-Axes', 29)
      call boxed(u, ' Atoms           Fractional Coordinate           Cartesian Coordinate (Normalized) ')
      write(u, '(1x,a)') plus
      do a = 1, inp%nb
        write(u, '(1x,a1,2x,a3,5x,3f10.5,5x,a1,2x,3f10.5,5x,a1)') '|', capital(inp%label(a)), &
          nz(frac_of(inp, inp%op_axes(:, a))), '|', nz(inp%op_axes(:, a)), '|'
      end do
      write(u, '(1x,a)') rule
      write(u, '(a)') '  '
    end if
    if (inp%lstaggered) then
      call title(u, 'Staggered Magnetization', 29)
      call boxed(u, ' Atom     Type          Fractional Coordinate           Staggered  Coefficient     ')
      write(u, '(1x,a)') plus
      do a = 1, inp%nb
        if (inp%stag_coeff(a) > 0) then
          write(u, '(1x,a1,1x,a3,6x,i3,7x,3f10.5,5x,a1,12x,a1,i1,13x,a1)') '|', capital(inp%label(a)), &
            inp%kind(a), inp%basis(:, a), '|', '+', nint(inp%stag_coeff(a)), '|'
        else
          write(u, '(1x,a1,1x,a3,6x,i3,7x,3f10.5,5x,a1,12x,i2,13x,a1)') '|', capital(inp%label(a)), &
            inp%kind(a), inp%basis(:, a), '|', nint(inp%stag_coeff(a)), '|'
        end if
      end do
      write(u, '(1x,a)') rule
      write(u, '(a)') '  '
    end if
    if (inp%lsfactor) then
      call title(u, 'STRUCTURE-FACTOR', 32)
      write(u, '(1x,a48,10x,i9,17x,a1)') '|   Number of steps to take a SF measurement  : ', inp%sf_steps_measure, '|'
      call boxed(u, '   2D q parameters (in reduced  coordinates):')
      v = reduced_of(inp, inp%sf_corner) / inp%lenconfac
      write(u, '(1x,a14,2x,3f10.5,38x,a1)') '|     Corner: ', v, '|'
      v = reduced_of(inp, inp%sf_q1) / inp%lenconfac
      write(u, '(1x,a14,2x,3f10.5,7x,a12,2x,i4,13x,a1)') '|    Vector1: ', v, ' Divisions:', inp%sf_mesh(1), '|'
      v = reduced_of(inp, inp%sf_q2) / inp%lenconfac
      write(u, '(1x,a14,2x,3f10.5,7x,a12,2x,i4,13x,a1)') '|    Vector2: ', v, ' Divisions:', inp%sf_mesh(2), '|'
      if (inp%lenconfac == 1.0_dp) then
        call boxed(u, '   2D q parameters (in cartesian coordinates) (Ang^-1):')
      else
        call boxed(u, '   2D q parameters (in cartesian coordinates) (Bohr^-1):')
      end if
      write(u, '(1x,a14,2x,3f10.5,38x,a1)') '|     Corner: ', inp%sf_corner, '|'
      write(u, '(1x,a14,2x,3f10.5,7x,a12,2x,i4,13x,a1)') '|    Vector1: ', inp%sf_q1, ' Divisions:', inp%sf_mesh(1), '|'
      write(u, '(1x,a14,2x,3f10.5,7x,a12,2x,i4,13x,a1)') '|    Vector2: ', inp%sf_q2, ' Divisions:', inp%sf_mesh(2), '|'
      write(u, '(1x,a)') rule
      write(u, '(a)') '  '
    end if
  end subroutine

  ! The random-number box (param_write_seed): the host MT19937 stream's seed (one rank here) and,
  ! with the disorder, the seed of its normal numbers.
  subroutine write_seed_box(u, inp, seed, sg_seed)
    integer, intent(in) :: u, seed, sg_seed
    type(input_t), intent(in) :: inp
    call title(u, 'Random Number Generator', 28)
    call boxed(u, '  Random number generatorg algorithm         :          Mersenne Twister           ')
    if (inp%have_seed) then
      call boxed(u, '  Initialized with seed derived from         :               *.mcin file           ')
    else
      call boxed(u, '  Initialized with seed derived from         :               Timer                 ')
    end if
    call boxed(u, '  Seeds:')
    write(u, '(1x,a1,5x,1i12,66x,a1)') '|', seed, '|'
    if (inp%spin_glass) then
      if (inp%have_sg_seed) then
        call boxed(u, '  Initialized spin glass seed                :  spin_glass_seed from *.mcin file   ')
      else
        call boxed(u, '  Initialized spin glass seed                :               Timer                 ')
      end if
      write(u, '(1x,a48,10x,i12,14x,a1)') '|  Used seed for normal random generator      : ', sg_seed, '|'
    end if
    write(u, '(1x,a)') rule
  end subroutine

end module mc_echo
