! mc_rng.f90 -- the host random stream of the driver: MT19937 (Matsumoto & Nishimura 1998),
! seeded like the reference's rank stream (comms_init_random, comms.F90:979-1007; seeds
! keyword, parameters.F90:388-398) so that the 'rand' initial configuration
! (mcarlo.F90:275-286, pt.F90:448-459) is the reference's for the same non-negative seed.
! Values u = u32 / 2^32 in [0,1), as mtprng_rand_real2 (mtprng.f90:347-361).
!
! The generator is written from the published algorithm: state recurrence
! x(i) = 1812433253 (x(i-1) xor (x(i-1) >> 30)) + i, a twist of 624 words with the matrix
! 0x9908B0DF, and tempering with the masks 0x9D2C5680 / 0xEFC60000.  Arithmetic is in 64-bit
! integers masked to 32 bits.
module mc_rng
  use iso_c_binding, only: c_int64_t
  implicit none
  private

  integer, parameter :: nn = 624, mm = 397
  integer(c_int64_t), parameter :: mask32 = int(z'FFFFFFFF', c_int64_t)
  integer(c_int64_t), parameter :: upper = int(z'80000000', c_int64_t)
  integer(c_int64_t), parameter :: lower = int(z'7FFFFFFF', c_int64_t)
  integer(c_int64_t), parameter :: matrix_a = int(z'9908B0DF', c_int64_t)

  type, public :: mt_t
    integer(c_int64_t) :: x(0:nn-1) = 0
    integer :: i = nn + 1
  end type

  public :: mt_seed, mt_u32, mt_real2

contains

  subroutine mt_seed(st, seed)
    type(mt_t), intent(out) :: st
    integer, intent(in) :: seed
    integer :: k
    st%x(0) = iand(int(seed, c_int64_t), mask32)
    do k = 1, nn - 1
      st%x(k) = iand(1812433253_c_int64_t * ieor(st%x(k-1), ishft(st%x(k-1), -30)) + k, mask32)
    end do
    st%i = nn
  end subroutine

  integer(c_int64_t) function mt_u32(st) result(y)
    type(mt_t), intent(inout) :: st
    integer :: k
    integer(c_int64_t) :: v
    if (st%i >= nn) then
      do k = 0, nn - 1
        v = ior(iand(st%x(k), upper), iand(st%x(mod(k + 1, nn)), lower))
        st%x(k) = ieor(st%x(mod(k + mm, nn)), ishft(v, -1))
        if (iand(v, 1_c_int64_t) /= 0) st%x(k) = ieor(st%x(k), matrix_a)
      end do
      st%i = 0
    end if
    y = st%x(st%i)
    st%i = st%i + 1
    y = ieor(y, ishft(y, -11))
    y = ieor(y, iand(ishft(y, 7), int(z'9D2C5680', c_int64_t)))
    y = ieor(y, iand(ishft(y, 15), int(z'EFC60000', c_int64_t)))
    y = ieor(y, ishft(y, -18))
    y = iand(y, mask32)
  end function

  double precision function mt_real2(st)
    type(mt_t), intent(inout) :: st
    mt_real2 = real(mt_u32(st), kind(1.0d0)) * 2.3283064365386962890625d-10
  end function

end module mc_rng
