! Direct-mode integral plug-in of hfsolver's RHF path.
!
! Module `gaussians`, a build-time alternative to src/gaussians_hfsolver.f90
! and src/gaussians_libint.f90 (chosen in src/CMakeLists.txt:67-77), used
! together with the direct `scf` module (scf_direct.f90). getints2 has the
! interface of src/gaussians_hfsolver.f90:13-34 and is called unchanged by
! gaussints (src/basis.f90:567). It
!   - creates the MI355X engine for the basis (Schwarz threshold 1e-12) and
!     keeps it in hfx_direct (module hfx_gpu), replacing the engine of any
!     earlier basis; the direct doscf builds G(P) with it every iteration;
!   - still fills int2 (store mode, every primitive pair kept) for the
!     consumers of the stored integrals (MBPT, Green's functions,
!     src/mbpt.f90, src/gf.f90).
! Errors stop with exit code 1, as stop_error does (src/utils.f90:171-186).
module gaussians
use iso_c_binding, only: c_double, c_associated, c_null_ptr
use hfx_gpu, only: hfx_init, hfx_destroy, hfx_store_int2, hfx_direct
implicit none
private
public getints2

integer, parameter :: dp = c_double

contains

subroutine getints2(nprim, istart, center, power, coef, alpha, int2)
integer, intent(in) :: nprim(:), istart(:)
real(dp), intent(in) :: center(:, :)
integer, intent(in) :: power(:, :)
real(dp), intent(in) :: coef(:), alpha(:)
real(dp), intent(out) :: int2(:)
real(dp), allocatable :: atoms(:, :)
integer, allocatable :: atno(:)
integer :: n, i, k, nat

n = size(nprim)
allocate(atoms(3, n))
! the engine groups functions by atom: the distinct centres stand for the
! nuclei (charges 0: only the two-electron integrals are asked for)
nat = 0
do i = 1, n
    do k = 1, nat
        if (all(atoms(:, k) == center(:, i))) exit
    end do
    if (k > nat) then
        nat = nat + 1
        atoms(:, nat) = center(:, i)
    end if
end do
allocate(atno(nat))
atno = 0
if (c_associated(hfx_direct)) then
    call hfx_destroy(hfx_direct)
    hfx_direct = c_null_ptr
end if
call hfx_init(nprim, istart, center, power, coef, alpha, atno, atoms(:, 1:nat), hfx_direct)
call hfx_store_int2(nprim, istart, center, power, coef, alpha, atno, atoms(:, 1:nat), int2)
end subroutine

end module
