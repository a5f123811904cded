! Direct-mode SCF of hfsolver's RHF path (module `scf`).
!
! doscf has the interface of src/scf.f90:72-136 and is called unchanged by
! rhf_gauss (src/drivers.f90:37). The two-electron matrix of every iteration,
! G(P) = J - K/2, comes from the MI355X engine that the direct getints2
! (gaussians_direct.f90) keeps in hfx_direct: one hfx_fock_2e call
! (hfx_formG) in place of G = formG(P, TT) (src/scf.f90:109), so the n^4
! tensor TT (src/scf.f90:79,84-105) is never built and int2 is not read.
! The iteration keeps the reference's semantics: P = 0 at the start, F = H + G,
! E0 from electronic_energy, F C = S C lam by eigh (src/linalg.f90:241-280),
! P = 2 C_occ C_occ^T, dP = sqrt(sum (P - Pold)^2 / 4), converged after more
! than three iterations once dP < tolP and the last four E0 lie within tolE,
! linear mixing P = (1 - alpha) Pold + alpha P otherwise, stop_error when Nscf
! iterations do not converge.
!
! Every other public name of the reference's scf module comes from
! src/scf.f90 itself, compiled unchanged under the module name hfsolver_scf
! (amdflang -cpp -Dscf=hfsolver_scf), and is re-exported here, so the
! modules that use scf (basis, mbpt, gf, ...) link as they are.
module scf
use types, only: dp
use linalg, only: eigh
use utils, only: stop_error, str
use iso_c_binding, only: c_associated
use hfx_gpu, only: hfx_direct, hfx_formG
use hfsolver_scf, only: electronic_energy, get_nuclear_energy, ijkl2intindex, ijkl2intindex2, &
    kinetic_energy, create_intindex_sym4
implicit none
private
public doscf, electronic_energy, get_nuclear_energy, ijkl2intindex, ijkl2intindex2, kinetic_energy, &
    create_intindex_sym4

contains

subroutine doscf(H, int2, S, Noccupied, Nscf, tolE, tolP, alpha, C, P, lam, E0)
real(dp), intent(in) :: H(:, :), int2(:), S(:, :)
integer, intent(in) :: Noccupied, Nscf
real(dp), intent(in) :: tolE, tolP, alpha
real(dp), intent(out) :: C(:, :), P(:, :), lam(:), E0
real(dp), dimension(size(H, 1), size(H, 1)) :: G, F, Pold
real(dp) :: Ehist(Nscf), drms, spread
integer :: it
if (.not. c_associated(hfx_direct)) &
    call stop_error("doscf: no direct engine (getints2 of the direct gaussians module creates it)")
print *, "doscf: G(P) from the direct J/K engine (hfx_fock_2e) every iteration"
P = 0
do it = 1, Nscf
    call hfx_formG(hfx_direct, P, G)
    F = H + G
    E0 = electronic_energy(P, H, F)
    Ehist(it) = E0
    print *, it, E0
    call eigh(F, S, lam, C)
    Pold = P
    P = 2 * matmul(C(:, :Noccupied), transpose(C(:, :Noccupied)))
    drms = sqrt(sum((P - Pold)**2) / 4)
    if (it > 3) then
        spread = maxval(Ehist(it-3:it)) - minval(Ehist(it-3:it))
        if (drms < tolP .and. spread < tolE) then
            print *, "Converged in", it, "iterations: dP =", drms, " E0 spread =", spread
            return
        end if
    end if
    P = (1 - alpha) * Pold + alpha * P
end do
call stop_error("SCF did not converge in " // str(Nscf) // " iterations.")
end subroutine

end module
