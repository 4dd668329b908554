! mc_input.f90 -- reader for the subset of the ESpinS .mcin format the hot path needs.
!
! Line syntax (parameters.F90:1724-1805): case-insensitive "key = v", "key : v" or "key v";
! '!' and '#' start comments; "begin X ... end X" blocks with an optional unit line.
! Atoms are sorted by species in order of first appearance (parameters.F90:2470-2540).
! Coupling rows "f1=x,y,z:f2=x,y,z:jij=v" or c1=/c2= Cartesian (parameters.F90:2558-2846).
module mc_input
  use iso_c_binding, only: c_double, c_int
  implicit none
  private

  integer, parameter, public :: dp = c_double
  real(dp), parameter, public :: k_B = 8.6173303e-5_dp          ! constants.F90:48
  real(dp), parameter, public :: hart = 27.21138602_dp          ! constants.F90:49
  real(dp), parameter, public :: bohr_magn = 5.7883818012e-5_dp ! constants.F90:50
  real(dp), parameter, public :: bohr = 0.5291772108_dp       ! constants.F90:37
  integer, parameter :: maxl = 512

  type, public :: rows_t
    integer :: n = 0
    real(dp), allocatable :: f1(:,:), f2(:,:), val(:), dvec(:,:)
    real(dp), allocatable :: dvec_frac(:,:)     ! DM vectors in fractional coordinates, as given or
                                                ! from the unnormalised Cartesian input (parameters.F90:868-893)
    integer, allocatable :: shell(:)            ! sh= of each row (0 when absent)
    real(dp), allocatable :: sigma(:)           ! sig= of each row (eV, 0 when absent)
  end type

  type, public :: input_t
    real(dp) :: lat(3,3)                        ! lat(:,k) = lattice vector k (Angstrom)
    integer :: nb = 0
    character(len=8), allocatable :: label(:)
    real(dp), allocatable :: basis(:,:), moment(:)
    integer :: L(3) = 4
    type(rows_t) :: J, B, D
    logical :: have_B = .false., have_D = .false., have_si = .false., have_h = .false.
    real(dp), allocatable :: si_par(:), si_axis(:,:), h_par(:), h_axis(:,:)
    character(len=16) :: tems_mode = 'lin', mcarlo_mode = 'random', initial = 'ferro'
    integer :: tems_num = 1, steps_warmup = 100000, steps_mc = 200000, steps_measure = 2
    ! mcarlo_mode = learned (an addition of this driver): records and epochs of the training
    ! between the warm-up (plain 'random' proposals) and the measured phase
    integer :: learned_records = 1048576, learned_epochs = 4
    real(dp) :: tem_start = 5.0_dp, tem_end = 5.0_dp, tilt = 0.125_dp
    real(dp), allocatable :: tems(:)
    logical :: pt = .false.
    integer :: pt_steps_swap = 1, pt_print_swap = 10000
    integer :: seed = 0
    logical :: have_seed = .false., open_boundary = .false.
    logical :: energy_write = .false.     ! _energy-TNNN.dat series (parameters.F90:536-537)
    ! units of the parameter echo only (parameters.F90:224-241): coupling values are converted to
    ! eV on input, lengths to Angstrom
    real(dp) :: paramconfac = 1.0_dp, lenconfac = 1.0_dp
    logical :: coord_cart = .false.
    ! species (in order of first appearance) and the species of each sorted atom
    integer :: nsp = 0
    character(len=8), allocatable :: species(:)
    integer, allocatable :: kind(:)
    ! optional observables (parameters.F90:440-592)
    logical :: lorderparam = .false., lstaggered = .false., lspincorr = .false.
    logical :: lbinerror = .false., lsfactor = .false., lsf_polar = .false.
    real(dp), allocatable :: op_axes(:,:), stag_coeff(:)
    integer, allocatable :: binning_level(:)
    integer :: sf_mesh(2) = 50, sf_steps_measure = 100
    real(dp) :: sf_corner(3) = 0, sf_q1(3) = 0, sf_q2(3) = 0, sf_polar(3) = 0
    real(dp) :: recip(3,3) = 0                  ! recip(:,k) = reciprocal vector k (2 pi / Angstrom)
    ! spin-glass disorder of the exchange (parameters.F90:616-630, jij.F90:506-556)
    logical :: spin_glass = .false., have_sg_seed = .false.
    integer :: sg_seed = 0
  end type

  public :: read_input, input_error

  character(len=maxl), allocatable :: lines(:)
  logical, allocatable :: used(:)      ! lines consumed by a keyword or block (param_read, parameters.F90:999-1030)
  integer :: nlines

contains

  subroutine input_error(msg)
    character(len=*), intent(in) :: msg
    write(*, '(1x,a)') 'Error: '//trim(msg)
    error stop 1
  end subroutine

  pure function lower(s) result(t)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: k, c
    t = s
    do k = 1, len(s)
      c = iachar(s(k:k))
      if (c >= 65 .and. c <= 90) t(k:k) = achar(c + 32)
    end do
  end function

  subroutine load(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios, k, p
    character(len=maxl) :: buf
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) call input_error('cannot open '//trim(fname))
    nlines = 0
    do
      read(u, '(a)', iostat=ios) buf
      if (ios /= 0) exit
      nlines = nlines + 1
    end do
    allocate(lines(nlines), used(nlines))
    used = .false.
    rewind(u)
    do k = 1, nlines
      read(u, '(a)') buf
      p = scan(buf, '!#')
      if (p > 0) buf(p:) = ' '
      do p = 1, len(buf)
        if (iachar(buf(p:p)) == 9) buf(p:p) = ' '
      end do
      lines(k) = adjustl(lower(buf))
    end do
    close(u)
  end subroutine

  ! value string of keyword key, or found = .false.
  subroutine keyword(key, found, val)
    character(len=*), intent(in) :: key
    logical, intent(out) :: found
    character(len=*), intent(out) :: val
    integer :: k, n
    character(len=maxl) :: rest
    found = .false.
    val = ' '
    n = len_trim(key)
    do k = 1, nlines
      if (lines(k)(1:n) /= key(1:n)) cycle
      rest = lines(k)(n+1:)
      if (len_trim(rest) > 0 .and. index(' =:', rest(1:1)) == 0) cycle
      rest = adjustl(rest)
      if (rest(1:1) == '=' .or. rest(1:1) == ':') rest = adjustl(rest(2:))
      if (found) call input_error('keyword '//trim(key)//' found more than once')
      found = .true.
      used(k) = .true.
      val = rest
    end do
  end subroutine

  logical function is_true(v)
    character(len=*), intent(in) :: v
    character(len=maxl) :: w
    w = adjustl(v)
    if (w(1:1) == '.') w = w(2:)
    is_true = w(1:1) == 't'
  end function

  ! [first, last] body lines of block name (units line included), or found = .false.
  subroutine block(name, found, first, last)
    character(len=*), intent(in) :: name
    logical, intent(out) :: found
    integer, intent(out) :: first, last
    integer :: k
    character(len=maxl) :: w
    found = .false.
    first = 0
    last = -1
    do k = 1, nlines
      if (lines(k)(1:5) /= 'begin') cycle
      w = adjustl(lines(k)(6:))
      if (trim(w) /= trim(name)) cycle
      if (found) call input_error('block '//trim(name)//' found more than once')
      found = .true.
      first = k + 1
    end do
    if (.not. found) return
    do k = first, nlines
      if (lines(k)(1:3) == 'end') then
        last = k - 1
        used(first - 1:k) = .true.
        return
      end if
    end do
    call input_error('begin '//trim(name)//' without end')
  end subroutine

  ! reciprocal vectors recip(:,k) with lat(:,j) . recip(:,k) = 2 pi delta_jk
  subroutine recip_of(lat, recip)
    real(dp), intent(in) :: lat(3,3)
    real(dp), intent(out) :: recip(3,3)
    real(dp), parameter :: twopi = 6.283185307179586476925286766559_dp
    real(dp) :: e(3), f(3)
    integer :: k
    do k = 1, 3
      e = 0
      e(k) = 1
      call cart_to_frac(lat, e, f)        ! f = column k of inv(lat)
      recip(k, :) = twopi * f             ! recip = 2 pi inv(lat)^T
    end do
  end subroutine

  ! Integer lists such as "1-5", "1 3 5" or "2,4-6" (param_get_range_vector, parameters.F90:2245-2341)
  subroutine range_vector(w, out)
    character(len=*), intent(in) :: w
    integer, allocatable, intent(out) :: out(:)
    character(len=len(w)) :: t
    integer :: p, q, a, b, k, n, buf(1024), ios
    t = w
    do p = 1, len(t)
      if (t(p:p) == ',' .or. t(p:p) == ';') t(p:p) = ' '
      if (t(p:p) == ':') t(p:p) = '-'
    end do
    n = 0
    p = 1
    do
      do while (p <= len_trim(t))
        if (t(p:p) /= ' ') exit
        p = p + 1
      end do
      if (p > len_trim(t)) exit
      q = p
      do while (q <= len_trim(t))
        if (t(q:q) == ' ') exit
        q = q + 1
      end do
      k = index(t(p:q-1), '-')
      if (k > 1) then
        read(t(p:p+k-2), *, iostat=ios) a
        if (ios /= 0) call input_error('Error parsing range '//trim(w))
        read(t(p+k:q-1), *, iostat=ios) b
        if (ios /= 0) call input_error('Error parsing range '//trim(w))
        do k = min(a, b), max(a, b)
          n = n + 1
          buf(n) = k
        end do
      else
        read(t(p:q-1), *, iostat=ios) a
        if (ios /= 0) call input_error('Error parsing range '//trim(w))
        n = n + 1
        buf(n) = a
      end if
      p = q
    end do
    allocate(out(n))
    out = buf(1:n)
  end subroutine

  subroutine cart_to_frac(lat, c, f)
    real(dp), intent(in) :: lat(3,3), c(3)
    real(dp), intent(out) :: f(3)
    real(dp) :: a(3,3), det, inv(3,3)
    a = lat
    det = a(1,1)*(a(2,2)*a(3,3)-a(3,2)*a(2,3)) - a(1,2)*(a(2,1)*a(3,3)-a(2,3)*a(3,1)) &
        + a(1,3)*(a(2,1)*a(3,2)-a(2,2)*a(3,1))
    inv(1,1) = (a(2,2)*a(3,3)-a(3,2)*a(2,3))/det
    inv(1,2) = (a(1,3)*a(3,2)-a(1,2)*a(3,3))/det
    inv(1,3) = (a(1,2)*a(2,3)-a(1,3)*a(2,2))/det
    inv(2,1) = (a(2,3)*a(3,1)-a(2,1)*a(3,3))/det
    inv(2,2) = (a(1,1)*a(3,3)-a(1,3)*a(3,1))/det
    inv(2,3) = (a(2,1)*a(1,3)-a(1,1)*a(2,3))/det
    inv(3,1) = (a(2,1)*a(3,2)-a(3,1)*a(2,2))/det
    inv(3,2) = (a(3,1)*a(1,2)-a(1,1)*a(3,2))/det
    inv(3,3) = (a(1,1)*a(2,2)-a(2,1)*a(1,2))/det
    f = matmul(inv, c)
  end subroutine

  ! one coupling block: rows in eV (after the ryd conversion)
  subroutine read_rows(inp, name, key, r)
    type(input_t), intent(in) :: inp
    character(len=*), intent(in) :: name, key
    type(rows_t), intent(out) :: r
    logical :: found
    integer :: first, last, k, n, p, q, ios
    real(dp) :: econv, lconv, v3(3)
    character(len=maxl) :: s, tok(8)
    integer :: nt
    call block(name, found, first, last)
    if (.not. found) return
    econv = 1.0_dp
    lconv = 1.0_dp
    if (first <= last .and. index(lines(first), 'f1=') == 0 .and. index(lines(first), 'c1=') == 0) then
      if (index(lines(first), 'ryd') > 0) econv = hart * 0.5_dp
      if (index(lines(first), 'bohr') > 0) lconv = bohr
      first = first + 1
    end if
    n = 0
    do k = first, last
      if (len_trim(lines(k)) > 0) n = n + 1
    end do
    r%n = n
    allocate(r%f1(3,n), r%f2(3,n), r%val(n), r%shell(n), r%sigma(n))
    r%shell = 0
    r%sigma = 0
    n = 0
    do k = first, last
      if (len_trim(lines(k)) == 0) cycle
      n = n + 1
      ! remove blanks, split on ':'
      s = ' '
      q = 0
      do p = 1, len_trim(lines(k))
        if (lines(k)(p:p) /= ' ') then
          q = q + 1
          s(q:q) = lines(k)(p:p)
        end if
      end do
      nt = 0
      p = 1
      do while (p <= len_trim(s) .and. nt < 8)
        q = index(s(p:), ':')
        nt = nt + 1
        if (q == 0) then
          tok(nt) = s(p:)
          exit
        end if
        tok(nt) = s(p:p+q-2)
        p = p + q
      end do
      if (nt < 3) call input_error('malformed coupling row: '//trim(lines(k)))
      call site(tok(1), '1', r%f1(:,n))
      call site(tok(2), '2', r%f2(:,n))
      if (tok(3)(1:len(key)+1) /= key//'=') call input_error('malformed parameter: '//trim(lines(k)))
      read(tok(3)(len(key)+2:), *, iostat=ios) r%val(n)
      if (ios /= 0) call input_error('bad value in '//trim(lines(k)))
      r%val(n) = r%val(n) * econv
      do p = 4, nt                ! the optional sh= field (param_get_parameters, parameters.F90:2751-2766)
        if (tok(p)(1:3) == 'sh=') then
          read(tok(p)(4:), *, iostat=ios) r%shell(n)
          if (ios /= 0) call input_error('bad sh= in '//trim(lines(k)))
        else if (tok(p)(1:4) == 'sig=') then
          read(tok(p)(5:), *, iostat=ios) r%sigma(n)
          if (ios /= 0) call input_error('bad sig= in '//trim(lines(k)))
          r%sigma(n) = r%sigma(n) * econv
        end if
      end do
    end do
  contains
    subroutine site(t, tag, f)
      character(len=*), intent(in) :: t, tag
      real(dp), intent(out) :: f(3)
      integer :: ios2
      if (t(1:3) == 'f'//tag//'=') then
        read(t(4:), *, iostat=ios2) f
      else if (t(1:3) == 'c'//tag//'=') then
        read(t(4:), *, iostat=ios2) v3
        call cart_to_frac(inp%lat, v3 * lconv, f)
      else
        call input_error('atom site not recognised: '//trim(t))
      end if
      if (ios2 /= 0) call input_error('bad position: '//trim(t))
    end subroutine
  end subroutine

  ! single-ion / field axes: "label ax ay az value" lines, sorted to the species order
  subroutine read_axes(inp, name, par, ax, found)
    type(input_t), intent(in) :: inp
    character(len=*), intent(in) :: name
    real(dp), allocatable, intent(out) :: par(:), ax(:,:)
    logical, intent(out) :: found
    logical :: fc, ff
    integer :: first, last, fc1, fc2, k, a, used(inp%nb), ios
    character(len=8) :: lab
    real(dp) :: v(3), p
    call block(name//'_cart', fc, fc1, fc2)
    call block(name//'_frac', ff, first, last)
    found = fc .or. ff
    if (.not. found) return
    if (fc) then
      first = fc1
      last = fc2
    end if
    allocate(par(inp%nb), ax(3, inp%nb))
    par = 0
    ax = 0
    used = 0
    do k = first, last
      if (len_trim(lines(k)) == 0) cycle
      if (lines(k)(1:3) == 'ang' .or. lines(k)(1:4) == 'bohr') cycle
      read(lines(k), *, iostat=ios) lab, v, p
      if (ios /= 0) call input_error('bad line in '//name//' block: '//trim(lines(k)))
      if (ff) v = matmul(inp%lat, v)
      v = v / sqrt(sum(v * v))
      do a = 1, inp%nb
        if (trim(inp%label(a)) == trim(lab) .and. used(a) == 0) then
          used(a) = 1
          ax(:, a) = v
          par(a) = p
          exit
        end if
        if (a == inp%nb) call input_error('too many lines for '//trim(lab)//' in '//name)
      end do
    end do
  end subroutine

  subroutine read_input(fname, inp)
    character(len=*), intent(in) :: fname
    type(input_t), intent(out) :: inp
    logical :: found, frac
    integer :: first, last, k, n, a, s, ios, nsp
    character(len=maxl) :: v
    character(len=8), allocatable :: lab(:), species(:)
    real(dp), allocatable :: pos(:,:), mom(:)
    integer, allocatable :: orig(:)
    real(dp) :: scale, step, x(3)
    real(dp), allocatable :: dv(:,:)

    call load(fname)

    ! lattice (rows are vectors, parameters.F90:414-421)
    call block('unit_cell_cart', found, first, last)
    if (.not. found) call input_error('Did not find the cell information in the input file')
    scale = 1.0_dp
    if (lines(first)(1:4) == 'bohr') then
      scale = bohr
      first = first + 1
    else if (lines(first)(1:3) == 'ang') then
      first = first + 1
    end if
    if (last - first + 1 /= 3) call input_error('Not correct the cell information in the input file')
    do k = 1, 3
      read(lines(first + k - 1), *, iostat=ios) inp%lat(:, k)
      if (ios /= 0) call input_error('bad unit_cell_cart line')
    end do
    inp%lat = inp%lat * scale

    ! atoms
    call block('atoms_frac', frac, first, last)
    if (.not. frac) then
      call block('atoms_cart', found, first, last)
      if (.not. found) call input_error('No atoms_frac or atoms_cart block found')
    end if
    scale = 1.0_dp
    if (lines(first)(1:4) == 'bohr') then
      scale = bohr
      first = first + 1
    else if (lines(first)(1:3) == 'ang') then
      first = first + 1
    end if
    n = last - first + 1
    allocate(lab(n), pos(3, n), mom(n), species(n))
    do k = 1, n
      read(lines(first + k - 1), *, iostat=ios) lab(k), pos(:, k), mom(k)
      if (ios /= 0) call input_error('bad atoms line: '//trim(lines(first + k - 1)))
      if (.not. frac) then
        call cart_to_frac(inp%lat, pos(:, k) * scale, x)
        pos(:, k) = x
      end if
      do a = 1, 3
        if (abs(pos(a, k) - nint(pos(a, k))) < 1e-5_dp) pos(a, k) = nint(pos(a, k))
      end do
      pos(:, k) = pos(:, k) - floor(pos(:, k))
    end do
    nsp = 0
    do k = 1, n
      if (all(species(1:nsp) /= lab(k))) then
        nsp = nsp + 1
        species(nsp) = lab(k)
      end if
    end do
    inp%nb = n
    allocate(inp%label(n), inp%basis(3, n), inp%moment(n), inp%kind(n), orig(n))
    a = 0
    do s = 1, nsp
      do k = 1, n
        if (lab(k) == species(s)) then
          a = a + 1
          inp%label(a) = lab(k)
          inp%basis(:, a) = pos(:, k)
          inp%moment(a) = mom(k)
          inp%kind(a) = s
          orig(a) = k
        end if
      end do
    end do
    inp%nsp = nsp
    allocate(inp%species(nsp))
    inp%species = species(1:nsp)
    call recip_of(inp%lat, inp%recip)

    call keyword('supercell_size', found, v)
    if (found) read(v, *) inp%L
    if (any(inp%L < 1)) call input_error('supercell_size must be greater than zero')

    call keyword('tems_mode', found, v)
    if (found) inp%tems_mode = v
    call keyword('tem_start', found, v)
    if (found) read(v, *) inp%tem_start
    call keyword('tem_end', found, v)
    if (found) read(v, *) inp%tem_end
    call keyword('tems_num', found, v)
    if (found) read(v, *) inp%tems_num
    if (inp%tems_num < 1) call input_error('tems_num must be greater than zero')
    allocate(inp%tems(inp%tems_num))
    ! temperature ladders, parameters.F90:293-326
    if (inp%tems_mode(1:3) == 'man') then
      call keyword('tems', found, v)
      if (.not. found) call input_error('tems_mode set to manual but no tems found')
      read(v, *, iostat=ios) inp%tems
      if (ios /= 0) call input_error('cannot read tems')
    else
      inp%tems(1) = inp%tem_start
      if (inp%tems_num > 1) then
        select case (inp%tems_mode(1:3))
        case ('lin')
          step = (inp%tem_end - inp%tem_start) / real(inp%tems_num - 1, dp)
          do k = 1, inp%tems_num
            inp%tems(k) = inp%tem_start + (k - 1) * step
          end do
        case ('inv')
          step = (1.0_dp / inp%tem_end - 1.0_dp / inp%tem_start) / real(inp%tems_num - 1, dp)
          do k = 1, inp%tems_num
            inp%tems(k) = 1.0_dp / (1.0_dp / inp%tem_start + (k - 1) * step)
          end do
        case ('log')
          do k = 1, inp%tems_num
            inp%tems(k) = inp%tem_start * exp(log(inp%tem_end / inp%tem_start) * &
                          (1.0_dp * (k - 1) / (inp%tems_num - 1)))
          end do
        case default
          call input_error('tems_mode not recognised')
        end select
      end if
    end if
    if (any(inp%tems <= 0)) call input_error('tems must be greater than zero')

    call keyword('steps_warmup', found, v)
    if (found) read(v, *) inp%steps_warmup
    if (inp%steps_warmup < 0) inp%steps_warmup = 0
    call keyword('steps_mc', found, v)
    if (found) read(v, *) inp%steps_mc
    if (inp%steps_mc < 1) call input_error('steps_mc must be greater than zero')
    call keyword('steps_measure', found, v)
    if (found) read(v, *) inp%steps_measure
    if (inp%steps_measure < 1) call input_error('steps_measure must be greater than zero')
    call keyword('initial_sconfig', found, v)
    if (found) inp%initial = v
    if (inp%initial(1:5) /= 'ferro' .and. inp%initial(1:4) /= 'rand' .and. inp%initial(1:4) /= 'file') &
      call input_error('Initial_sconfig not recognised')
    call keyword('mcarlo_mode', found, v)
    if (found) inp%mcarlo_mode = v
    if (inp%mcarlo_mode(1:5) /= 'const' .and. inp%mcarlo_mode(1:4) /= 'rand' .and. &
        inp%mcarlo_mode(1:5) /= 'learn') &
      call input_error('mcarlo_mode not recognised')
    call keyword('energy_write', found, v)
    if (found) inp%energy_write = is_true(v)
    call keyword('learned_records', found, v)
    if (found) read(v, *) inp%learned_records
    if (inp%learned_records < 1) call input_error('learned_records must be greater than zero')
    call keyword('learned_epochs', found, v)
    if (found) read(v, *) inp%learned_epochs
    if (inp%learned_epochs < 1) call input_error('learned_epochs must be greater than zero')
    call keyword('tilt_angles_max', found, v)
    if (found) read(v, *) inp%tilt
    call keyword('seeds', found, v)
    if (found) then
      read(v, *) inp%seed
      inp%have_seed = .true.
    end if
    call keyword('pt', found, v)
    if (found) inp%pt = is_true(v)
    if (inp%pt .and. inp%tems_num <= 1) call input_error('tems_num must be greater than one in parallel tempering')
    call keyword('pt_steps_swap', found, v)
    if (found) read(v, *) inp%pt_steps_swap
    if (inp%pt .and. inp%pt_steps_swap < 1) call input_error('pt_steps_swap must be greater than zero')
    call keyword('pt_print_swap', found, v)
    if (found) read(v, *) inp%pt_print_swap
    call keyword('boundary', found, v)
    if (found) inp%open_boundary = v(1:4) == 'open'
    if (inp%pt) call sort(inp%tems)

    ! Hamiltonian
    call read_rows(inp, 'jij_parameters', 'jij', inp%J)
    call keyword('ham_bij', found, v)
    if (found) inp%have_B = is_true(v)
    call keyword('ham_dij', found, v)
    if (found) inp%have_D = is_true(v)
    if (inp%have_B) call read_rows(inp, 'bij_parameters', 'bij', inp%B)
    if (inp%have_D) then
      call read_rows(inp, 'dij_parameters', 'dij', inp%D)
      call block('dij_vectors_cart', found, first, last)
      frac = .false.
      if (.not. found) then
        call block('dij_vectors_frac', found, first, last)
        frac = found
      end if
      if (.not. found) call input_error('no dij_vectors_frac or dij_vectors_cart block found')
      if (lines(first)(1:3) == 'ang' .or. lines(first)(1:4) == 'bohr') first = first + 1
      if (last - first + 1 /= inp%D%n) call input_error('Wrong number of lines in dij_vectors block')
      allocate(dv(3, inp%D%n), inp%D%dvec_frac(3, inp%D%n))
      do k = 1, inp%D%n
        read(lines(first + k - 1), *) dv(:, k)
        if (frac) then
          inp%D%dvec_frac(:, k) = dv(:, k)
          dv(:, k) = matmul(inp%lat, dv(:, k))
        else
          inp%D%dvec_frac(:, k) = matmul(transpose(inp%recip), dv(:, k)) / (2.0_dp * acos(-1.0_dp))
        end if
        dv(:, k) = dv(:, k) / sqrt(sum(dv(:, k)**2))
      end do
      call move_alloc(dv, inp%D%dvec)
    end if
    call keyword('ham_singleion', found, v)
    if (found) inp%have_si = is_true(v)
    if (inp%have_si) then
      call read_axes(inp, 'singleion_axes', inp%si_par, inp%si_axis, found)
      if (.not. found) call input_error('ham_singleion is true but no singleion_axes block found')
    end if
    call keyword('ham_field', found, v)
    if (found) inp%have_h = is_true(v)
    if (inp%have_h) then
      call read_axes(inp, 'field_axes', inp%h_par, inp%h_axis, found)
      if (.not. found) call input_error('ham_field is true but no field_axes block found')
    end if
    call observable_keywords()
    call ignored_keywords()
    call unsupported_keywords()
    call unrecognised_keywords(fname)
  contains
    ! The optional observables (parameters.F90:440-592): flags, axes, coefficients, binning
    ! levels and the q-mesh of the structure factor, with the reference's checks.
    subroutine observable_keywords()
      character(len=maxl) :: w
      integer :: ic(inp%nb), q
      logical :: ff, fc
      real(dp) :: v3(3), cr(3)
      call keyword('staggered_m', found, w)
      if (found) inp%lstaggered = is_true(w)
      if (inp%lstaggered .and. inp%nb == 1) then
        write(*, '(1x,a)') 'WARNING: Calculation of staggered magnetization for one atom is'
        write(*, '(9x,a)') ' meaningless staggered_m was changed to .False.'
        inp%lstaggered = .false.
      end if
      allocate(inp%stag_coeff(inp%nb))
      inp%stag_coeff = 1
      call keyword('staggered_m_coeff', found, w)
      if (found) then
        read(w, *, iostat=ios) ic
        if (ios /= 0) call input_error('staggered_m_coeff needs one integer per atom')
        if (inp%lstaggered .and. any(abs(ic) /= 1)) call input_error('staggered_m_coeff must be +1 or -1')
        do q = 1, inp%nb                ! input atom order -> species-sorted order
          inp%stag_coeff(q) = ic(orig(q))
        end do
      end if
      call keyword('order_parameter', found, w)
      if (found) inp%lorderparam = is_true(w)
      call axes_block(ff, fc)
      if (ff .and. fc) call input_error('Cannot specify both order_parameter_axes_frac and order_parameter_axes_cart')
      if (inp%lorderparam .and. .not. (ff .or. fc)) &
        call input_error('No order_parameter_axes_frac or order_parameter_axes_cart block found')
      call keyword('binning_error', found, w)
      if (found) inp%lbinerror = is_true(w)
      call keyword('binning_level', found, w)
      if (found) then
        call range_vector(w, inp%binning_level)
        if (size(inp%binning_level) < 1) call input_error('problem reading binning_level')
      else
        allocate(inp%binning_level(5))
        inp%binning_level = [1, 2, 3, 4, 5]
      end if
      if (inp%lbinerror) then
        if (any(inp%binning_level <= 0) .or. any(inp%binning_level > 30)) &
          call input_error('binning_level contain non-valid numbers. It must be greater than zero and as ' &
                           //'(steps_mc/2^(binning_level))-1>0.')
        if (any(inp%steps_mc / (2**inp%binning_level) <= 1)) &
          call input_error('binning_level contain non-valid numbers. It must be greater than zero and as ' &
                           //'(steps_mc/2^(binning_level))-1>0.')
      end if
      call keyword('spin_correlation', found, w)
      if (found) inp%lspincorr = is_true(w)
      if (inp%lspincorr .and. any(inp%J%shell <= 0)) &
        call input_error('sh in jij_parameters block must be greater than zero')
      call keyword('sfactor', found, w)
      if (found) inp%lsfactor = is_true(w)
      v3 = [1.0_dp, 0.0_dp, 0.0_dp]
      call keyword('sfactor_polar', found, w)
      if (found) then
        read(w, *, iostat=ios) v3
        if (ios /= 0) call input_error('bad sfactor_polar')
        inp%lsf_polar = .true.
      end if
      inp%sf_polar = matmul(inp%recip, v3)
      v3 = 0
      call keyword('sfactor_corner', found, w)
      if (found) read(w, *, iostat=ios) v3
      inp%sf_corner = matmul(inp%recip, v3)
      v3 = [1.0_dp, 0.0_dp, 0.0_dp]
      call keyword('sfactor_q1', found, w)
      if (found) read(w, *, iostat=ios) v3
      inp%sf_q1 = matmul(inp%recip, v3)
      v3 = [0.0_dp, 0.0_dp, 1.0_dp]
      call keyword('sfactor_q2', found, w)
      if (found) read(w, *, iostat=ios) v3
      inp%sf_q2 = matmul(inp%recip, v3)
      cr = [inp%sf_q1(2)*inp%sf_q2(3) - inp%sf_q1(3)*inp%sf_q2(2), inp%sf_q1(3)*inp%sf_q2(1) - inp%sf_q1(1)*inp%sf_q2(3), &
            inp%sf_q1(1)*inp%sf_q2(2) - inp%sf_q1(2)*inp%sf_q2(1)]
      if (sqrt(sum(cr**2)) < 1e-8_dp) call input_error('Vectors sfactor_q1 and sfactor_q2 not linearly independent')
      call keyword('sfactor_steps_measure', found, w)
      if (found) read(w, *, iostat=ios) inp%sf_steps_measure
      if (inp%lsfactor .and. (inp%sf_steps_measure < 1 .or. inp%sf_steps_measure > inp%steps_mc)) &
        call input_error('sfactor_steps_measure must be greater than zero and less than steps_mc.')
      call keyword('sfactor_2dqmesh', found, w)
      if (found) read(w, *, iostat=ios) inp%sf_mesh
      if (inp%lsfactor .and. any(inp%sf_mesh < 1)) call input_error('sfactor_2dqmesh must be greater than zero')
    end subroutine

    ! order_parameter_axes_frac / _cart: "label x y z" lines, sorted to the species order,
    ! normalised (param_get_axes with lsort and lnorm, parameters.F90:3061-3280)
    subroutine axes_block(ff, fc)
      logical, intent(out) :: ff, fc
      integer :: f1, f2, c1, c2, k, a, taken(inp%nb), cnt
      character(len=8) :: lab
      real(dp) :: v(3)
      call block('order_parameter_axes_frac', ff, f1, f2)
      call block('order_parameter_axes_cart', fc, c1, c2)
      if (.not. (ff .or. fc) .or. (ff .and. fc)) return
      if (fc) then
        f1 = c1
        f2 = c2
      end if
      if (lines(f1)(1:3) == 'ang' .or. lines(f1)(1:4) == 'bohr') f1 = f1 + 1
      allocate(inp%op_axes(3, inp%nb))
      taken = 0
      cnt = 0
      do k = f1, f2
        if (len_trim(lines(k)) == 0) cycle
        read(lines(k), *, iostat=ios) lab, v
        if (ios /= 0) call input_error('bad line in order_parameter_axes block: '//trim(lines(k)))
        if (ff) v = matmul(inp%lat, v)
        v = v / sqrt(sum(v * v))
        do a = 1, inp%nb
          if (trim(inp%label(a)) == trim(lab) .and. taken(a) == 0) then
            taken(a) = 1
            inp%op_axes(:, a) = v
            cnt = cnt + 1
            exit
          end if
        end do
      end do
      if (cnt /= inp%nb) &
        call input_error('Wrong number of lines in order_parameter_axes_frac or order_parameter_axes_cart block')
    end subroutine

    ! Reference keywords of features this driver does not run yet: accepted when off, an
    ! error when switched on (never silently ignored).
    subroutine unsupported_keywords()
      character(len=maxl) :: w
      call keyword('spin_glass', found, w)
      if (found) inp%spin_glass = is_true(w)
      call keyword('spin_glass_seed', found, w)
      if (found) then
        read(w, *, iostat=ios) inp%sg_seed
        if (ios /= 0) call input_error('bad spin_glass_seed')
        inp%have_sg_seed = .true.
      end if
      if (inp%spin_glass .and. any(inp%J%sigma < 0)) &
        call input_error('sig in jij_parameters block must be positive')
    end subroutine

    ! The units and coordinates of the parameter echo, and keywords of the reference that only
    ! its input-generation modes (-inp1/-inp2/-ham) read or that change only how its output is
    ! buffered: accepted with their reference checks.
    subroutine ignored_keywords()
      character(len=maxl) :: w
      call keyword('parameter_unit', found, w)
      if (found .and. w(1:2) /= 'ev' .and. w(1:12) /= 'electronvolt' .and. w(1:3) /= 'ryd') &
        call input_error('Value of parameter_unit not recognised in param_read')
      if (found .and. w(1:3) == 'ryd') inp%paramconfac = 2.0_dp / hart
      call keyword('coordinate', found, w)
      if (found .and. w(1:4) /= 'frac' .and. w(1:4) /= 'cart') &
        call input_error('Value of coordinate not recognised in param_read')
      if (found) inp%coord_cart = w(1:4) == 'cart'
      call keyword('length_unit', found, w)
      if (found .and. w(1:3) /= 'ang' .and. trim(w) /= 'bohr') &
        call input_error('Value of length_unit not recognised in param_read')
      if (found .and. trim(w) == 'bohr') inp%lenconfac = 1.0_dp / bohr
      call keyword('neighbors_tol', found, w)
      call keyword('inputfile_setup1', found, w)
      call keyword('inputfile_setup2', found, w)
      call keyword('hamfile_setup', found, w)
      call keyword('shells_jij', found, w)
      call keyword('shells_bij', found, w)
      call keyword('shells_dij', found, w)
      call keyword('energy_num_print', found, w)
      if (found) then
        read(w, *, iostat=ios) k
        if (ios /= 0 .or. k < 1) call input_error('energy_num_print must be greater than zero')
      end if
    end subroutine

    ! Any non-blank line no keyword or block consumed is fatal, as in the reference
    ! (parameters.F90:999-1030): the lines are printed, then the run stops.
    subroutine unrecognised_keywords(fn)
      character(len=*), intent(in) :: fn
      integer :: q
      if (all(used .or. len_trim(lines) == 0)) return
      write(*, '(1x,a)') 'The following section of file '//trim(fn)//' contained unrecognised keywords'
      write(*, *)
      do q = 1, nlines
        if (.not. used(q) .and. len_trim(lines(q)) > 0) write(*, '(1x,a)') trim(lines(q))
      end do
      write(*, *)
      call input_error('Unrecognised keyword(s) in input file')
    end subroutine
    subroutine sort(t)
      real(dp), intent(inout) :: t(:)
      integer :: i, j
      real(dp) :: tmp
      do i = 2, size(t)
        tmp = t(i)
        j = i - 1
        do while (j >= 1)
          if (t(j) <= tmp) exit
          t(j + 1) = t(j)
          j = j - 1
        end do
        t(j + 1) = tmp
      end do
    end subroutine
  end subroutine read_input

end module mc_input
