! Harness of ours around the reference's mtprng module (test infrastructure only).
! Usage: mt_ref SEED COUNT  -> prints COUNT outputs of mtprng_rand64, then the same
! stream through mtprng_rand_real2 in ES25.17, one per line.
program mt_ref
  use stdtypes
  use mtprng
  implicit none
  type(mtprng_state) :: st
  integer(INT32) :: seed
  integer :: n, i
  character(len=32) :: arg
  call get_command_argument(1, arg); read(arg, *) seed
  call get_command_argument(2, arg); read(arg, *) n
  call mtprng_init(seed, st)
  do i = 1, n
     write(*, '(I0)') mtprng_rand64(st)
  end do
  call mtprng_init(seed, st)
  do i = 1, n
     write(*, '(ES25.17)') mtprng_rand_real2(st)
  end do
end program mt_ref
