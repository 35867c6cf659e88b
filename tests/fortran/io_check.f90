!> Exercises the driver's file module (bdrun_io) on the CPU: .cfg in, then
!> dump / read back / dump again, a trajectory created, reopened and
!> appended, and a statistics file written and appended as after a revive.
!> Arguments: the .cfg, then the output directory.
program io_check
use iso_c_binding
use bdrun_io
implicit none
type(cfg_t) :: c, d
type(traj_t) :: t
integer(c_int64_t) :: nts, seed, offset
logical :: has
real(c_double) :: s(30)
integer :: fu, k
character(len=512) :: fn, od
call get_command_argument(1, fn)
call get_command_argument(2, od)
call read_config(trim(fn), c)
c%molecule_com = [0.5d0, -0.25d0, 1.0d-3]
call write_dump(trim(od)//'/a.dump', 1234_c_int64_t, c, 42_c_int64_t, 1000_c_int64_t)
call read_dump(trim(od)//'/a.dump', nts, d, has, seed, offset)
if (.not. has) stop 3
call write_dump(trim(od)//'/b.dump', nts, d, seed, offset)
call traj_create(trim(od)//'/t.bin', c%num_atoms, t)
call traj_append(t, 2_c_int64_t, c%coordinates)
call traj_append(t, 4_c_int64_t, 2 * c%coordinates)
call traj_close(t)
call traj_open(trim(od)//'/t.bin', t)
if (t%num_frames /= 2) stop 4
call traj_append(t, 6_c_int64_t, 3 * c%coordinates)
call traj_close(t)
s = [(k * 1.5d0 - 7.0d0, k = 1, 30)]
s(5) = 1.234567891d123
s(6) = -2.5d-201
call stats_open(trim(od)//'/s.txt', .false., c, .true., fu)
call stats_write_line(fu, 10_c_int64_t, c, .true., s)
close(fu)
call stats_open(trim(od)//'/s.txt', .true., c, .true., fu)
call stats_write_line(fu, 20_c_int64_t, c, .true., s * 1.0d-3)
close(fu)
end program
