! iso_c_binding interface to the MI355X direct J/K engine (include/hfx.h).
!
! Drop-in for hfsolver's closed-shell RHF path: a `doscf` built with this
! module replaces `G = formG(P, TT)` (src/scf.f90:109) by `call hfx_formG(h, P, G)`
! and skips the int2 / TT precompute (src/basis.f90:537-567,
! src/scf.f90:83-105). The interfaces follow the style of src/qc.f90:9-66
! (scalars by value, arrays by reference); a non-zero status is turned into
! stop_error semantics (message, exit code 1) as src/utils.f90:171-186 does.
module hfx_gpu
use iso_c_binding, only: c_int, c_double, c_long_long, c_ptr, c_null_ptr, c_char, c_associated, c_f_pointer
implicit none
private
public hfx_opts, hfx_default_opts, hfx_create, hfx_fock_2e, hfx_one_electron, hfx_destroy, &
    hfx_last_error, hfx_init, hfx_formG, hfx_store_int2, hfx_direct

! The engine of the current basis in direct mode: created (and replaced for
! every new basis) by getints2 of the direct `gaussians` module
! (gaussians_direct.f90), used by the direct `doscf` (scf_direct.f90).
type(c_ptr), save :: hfx_direct = c_null_ptr

type, bind(c) :: hfx_opts
    real(c_double) :: schwarz_threshold
    real(c_double) :: fp32_threshold
    integer(c_int) :: device
    integer(c_int) :: verbose
    integer(c_int) :: kernel_mode
end type

interface
    integer(c_int) function hfx_default_opts(o) bind(c, name="hfx_default_opts")
    import :: c_int, hfx_opts
    type(hfx_opts), intent(out) :: o
    end function

    integer(c_int) function hfx_create(nbf, nprim, istart, x, y, z, l, m, n, n2, coef, alpha, &
            nat, atno, xyz, o, h) bind(c, name="hfx_create")
    import :: c_int, c_double, c_ptr, hfx_opts
    integer(c_int), value, intent(in) :: nbf, n2, nat
    integer(c_int), intent(in) :: nprim(nbf), istart(nbf), l(nbf), m(nbf), n(nbf), atno(nat)
    real(c_double), intent(in) :: x(nbf), y(nbf), z(nbf), coef(n2), alpha(n2), xyz(3, nat)
    type(hfx_opts), intent(in) :: o
    type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function hfx_fock_2e(h, P, G, J, K, stats) bind(c, name="hfx_fock_2e")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value, intent(in) :: h
    real(c_double), intent(in) :: P(*)
    real(c_double), intent(out) :: G(*)
    type(c_ptr), value, intent(in) :: J, K, stats
    end function

    integer(c_int) function hfx_one_electron(h, S, T, V) bind(c, name="hfx_one_electron")
    import :: c_int, c_double, c_ptr
    type(c_ptr), value, intent(in) :: h
    real(c_double), intent(out) :: S(*), T(*), V(*)
    end function

    integer(c_int) function hfx_int2_c(h, int2, size) bind(c, name="hfx_int2")
    import :: c_int, c_double, c_long_long, c_ptr
    type(c_ptr), value, intent(in) :: h
    real(c_double), intent(out) :: int2(*)
    integer(c_long_long), value, intent(in) :: size
    end function

    type(c_ptr) function hfx_last_error_c(h) bind(c, name="hfx_last_error")
    import :: c_ptr
    type(c_ptr), value, intent(in) :: h
    end function

    subroutine hfx_destroy(h) bind(c, name="hfx_destroy")
    import :: c_ptr
    type(c_ptr), value, intent(in) :: h
    end subroutine
end interface

contains

function hfx_last_error(h) result(msg)
type(c_ptr), intent(in) :: h
character(len=512) :: msg
character(kind=c_char), pointer :: s(:)
integer :: i
msg = ""
if (.not. c_associated(hfx_last_error_c(h))) return
call c_f_pointer(hfx_last_error_c(h), s, [512])
do i = 1, 512
    if (s(i) == char(0)) exit
    msg(i:i) = s(i)
end do
end function

subroutine hfx_fail(h, where)
type(c_ptr), intent(in) :: h
character(*), intent(in) :: where
print "(a)", "ERROR STOP: " // where // ": " // trim(hfx_last_error(h))
stop 1
end subroutine

! Engine for the per-function basis arrays of gaussints (src/basis.f90:541-567);
! istart is 1-based here, as gaussints holds it, and passed 0-based as getInts does.
subroutine hfx_init(nprim, istart, center, power, coef, alpha, atno, xyz, h, fp32_threshold, schwarz_threshold)
integer, intent(in) :: nprim(:), istart(:), power(:, :), atno(:)
real(c_double), intent(in) :: center(:, :), coef(:), alpha(:), xyz(:, :)
type(c_ptr), intent(out) :: h
real(c_double), intent(in), optional :: fp32_threshold, schwarz_threshold
type(hfx_opts) :: o
integer(c_int) :: rc
rc = hfx_default_opts(o)
if (present(fp32_threshold)) o%fp32_threshold = fp32_threshold
if (present(schwarz_threshold)) o%schwarz_threshold = schwarz_threshold
rc = hfx_create(size(nprim), nprim, istart - 1, center(1, :), center(2, :), center(3, :), &
    power(1, :), power(2, :), power(3, :), size(coef), coef, alpha, size(atno), atno, xyz, o, h)
if (rc /= 0) call hfx_fail(h, "hfx_create")
end subroutine

! Every unique (ij|kl) in the packed order of ijkl2intindex (store mode,
! hfx_int2), from an engine of its own that keeps every primitive pair.
subroutine hfx_store_int2(nprim, istart, center, power, coef, alpha, atno, xyz, int2)
integer, intent(in) :: nprim(:), istart(:), power(:, :), atno(:)
real(c_double), intent(in) :: center(:, :), coef(:), alpha(:), xyz(:, :)
real(c_double), intent(out) :: int2(:)
type(c_ptr) :: h
call hfx_init(nprim, istart, center, power, coef, alpha, atno, xyz, h, schwarz_threshold=0.0_c_double)
if (hfx_int2_c(h, int2, int(size(int2), c_long_long)) /= 0) call hfx_fail(h, "hfx_int2")
call hfx_destroy(h)
end subroutine

! G(P) = J - K/2: the drop-in for G = formG(P, TT) (src/scf.f90:26-37).
subroutine hfx_formG(h, P, G)
type(c_ptr), intent(in) :: h
real(c_double), intent(in) :: P(:, :)
real(c_double), intent(out) :: G(:, :)
if (hfx_fock_2e(h, P, G, c_null_ptr, c_null_ptr, c_null_ptr) /= 0) call hfx_fail(h, "hfx_fock_2e")
end subroutine

end module
