! Test-infrastructure driver for the compiled reference (oracle/_ref only).
!
! Runs hfsolver's closed-shell RHF exactly as rhf_gauss does
! (src/drivers.f90:29-41: gaussints -> H=T+V -> doscf -> Enuc, Ekin, Etot),
! but keeps the intermediate arrays so that they can be written out as golden
! fixtures: S, T, V, C, lambda, the output density P, the two-electron matrix
! G(P) formed from the reference's packed int2 with the formG2 loop semantics
! (src/scf.f90:39-58), and a list of packed ERIs.
!
! It also records the wall time of gaussints and of doscf (system_clock), so
! that the reference's CPU path can be timed beside the GPU engine.
!
! Usage: ref_driver <input> <output>
!   input:  nat / (Z x y z) * nat [bohr] / Nscf alpha tolE tolP / nint2_dump
!   output: plain text, one "key n" header followed by n values per block.
! The process must run with a working directory holding p631ss.txt
! (src/basis.f90:210).
program ref_driver
use types, only: dp
use basis, only: gaussints
use scf, only: doscf, get_nuclear_energy, kinetic_energy, ijkl2intindex
implicit none
integer :: nat, i, j, k, l, n, Nscf, u, nocc, ndump, m
integer, allocatable :: atno(:)
real(dp), allocatable :: xyz(:, :), C(:, :), lam(:), int2(:), S(:, :), &
    T(:, :), V(:, :), P(:, :), H(:, :), G(:, :), Jm(:, :), Km(:, :)
real(dp) :: alpha, tolE, tolP, E0, Enuc, Ekin, Etot, t_ints, t_scf
integer(8) :: c0, c1, c2, crate
character(len=1024) :: inpath, outpath

call get_command_argument(1, inpath)
call get_command_argument(2, outpath)
open(newunit=u, file=trim(inpath), status="old")
read(u, *) nat
allocate(atno(nat), xyz(3, nat))
do i = 1, nat
    read(u, *) atno(i), xyz(:, i)
end do
read(u, *) Nscf, alpha, tolE, tolP
read(u, *) ndump
close(u)

call system_clock(c0, crate)
call gaussints(atno, xyz, S, T, V, int2)
call system_clock(c1)
n = size(S, 1)
allocate(H(n, n), P(n, n), C(n, n), lam(n), G(n, n), Jm(n, n), Km(n, n))
H = T + V
nocc = sum(atno) / 2
call doscf(H, int2, S, nocc, Nscf, tolE, tolP, alpha, C, P, lam, E0)
call system_clock(c2)
t_ints = real(c1 - c0, dp) / real(crate, dp)
t_scf = real(c2 - c1, dp) / real(crate, dp)
Enuc = get_nuclear_energy(atno, xyz)
Ekin = kinetic_energy(P, T)
Etot = E0 + Enuc

! G(P) at the output density, with the reference's own packed integrals.
do i = 1, n
    do j = 1, n
        Jm(i, j) = 0
        Km(i, j) = 0
        do k = 1, n
            do l = 1, n
                Jm(i, j) = Jm(i, j) + P(k, l) * int2(ijkl2intindex(i, j, k, l))
                Km(i, j) = Km(i, j) + P(k, l) * int2(ijkl2intindex(i, l, k, j))
            end do
        end do
    end do
end do
G = Jm - 0.5_dp * Km

open(newunit=u, file=trim(outpath), status="replace")
write(u, '(a, i12)') "n", 1
write(u, '(i12)') n
call wscal("E0", E0)
call wscal("Enuc", Enuc)
call wscal("Ekin", Ekin)
call wscal("Etot", Etot)
! wall times: gaussints (S, T, V and the packed int2) and doscf (TT precompute
! plus every SCF iteration)
call wscal("t_gaussints", t_ints)
call wscal("t_doscf", t_scf)
call wvec("lam", lam)
call wvec("S", reshape(S, [n*n]))
call wvec("T", reshape(T, [n*n]))
call wvec("V", reshape(V, [n*n]))
call wvec("C", reshape(C, [n*n]))
call wvec("P", reshape(P, [n*n]))
call wvec("J", reshape(Jm, [n*n]))
call wvec("K", reshape(Km, [n*n]))
call wvec("G", reshape(G, [n*n]))
m = size(int2)
if (ndump > 0) then
    ! every stride-th packed integral, with its 1-based (i,j,k,l), i>=j, k>=l, ij>=kl
    ! last block: the count is "-1", entries run to the end of the file
    write(u, '(a, i12)') "int2 ", -1
    call dump_int2(max(1, m / ndump))
end if
close(u)

contains

subroutine wscal(key, x)
character(*), intent(in) :: key
real(dp), intent(in) :: x
write(u, '(a, a, i12)') key, " ", 1
write(u, '(es25.17)') x
end subroutine

subroutine wvec(key, x)
character(*), intent(in) :: key
real(dp), intent(in) :: x(:)
integer :: q
write(u, '(a, a, i12)') key, " ", size(x)
do q = 1, size(x)
    write(u, '(es25.17)') x(q)
end do
end subroutine

subroutine dump_int2(stride)
integer, intent(in) :: stride
integer :: ii, jj, kk, ll, idx, cnt
cnt = 0
do ii = 1, n
    do jj = 1, ii
        do kk = 1, n
            do ll = 1, kk
                if ((ii-1)*ii/2 + jj < (kk-1)*kk/2 + ll) cycle
                idx = ijkl2intindex(ii, jj, kk, ll)
                if (mod(idx - 1, stride) /= 0) cycle
                if (cnt >= ndump) return
                cnt = cnt + 1
                write(u, '(4i6, es25.17)') ii, jj, kk, ll, int2(idx)
            end do
        end do
    end do
end do
end subroutine

end program
