! Test program for the Fortran binding (hfsolver_amd/fortran/hfx_gpu.f90):
! reads the per-function basis arrays and a density written by
! tests/test_gpu_fortran.py, calls hfx_init + hfx_formG exactly as a modified
! doscf would, and writes G.
program hfx_fortran_check
use iso_c_binding, only: c_ptr, c_double
use hfx_gpu, only: hfx_init, hfx_formG, hfx_destroy
implicit none
integer :: nbf, n2, nat, i, u
integer, allocatable :: nprim(:), istart(:), power(:, :), atno(:)
real(c_double), allocatable :: center(:, :), coef(:), alpha(:), xyz(:, :), P(:, :), G(:, :)
type(c_ptr) :: h
character(len=1024) :: inp, outp
call get_command_argument(1, inp)
call get_command_argument(2, outp)
open(newunit=u, file=trim(inp), status="old")
read(u, *) nbf, n2, nat
allocate(nprim(nbf), istart(nbf), power(3, nbf), center(3, nbf), coef(n2), alpha(n2), atno(nat), &
    xyz(3, nat), P(nbf, nbf), G(nbf, nbf))
do i = 1, nbf
    read(u, *) nprim(i), istart(i), power(:, i), center(:, i)
end do
do i = 1, n2
    read(u, *) coef(i), alpha(i)
end do
do i = 1, nat
    read(u, *) atno(i), xyz(:, i)
end do
read(u, *) P
close(u)
call hfx_init(nprim, istart, center, power, coef, alpha, atno, xyz, h)
call hfx_formG(h, P, G)
call hfx_destroy(h)
open(newunit=u, file=trim(outp), status="replace")
write(u, "(es25.17)") G
close(u)
end program
