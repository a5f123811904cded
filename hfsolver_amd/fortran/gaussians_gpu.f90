! GPU implementation of hfsolver's two-electron integral plug-in.
!
! The reference chooses its `gaussians` module at build time
! (src/CMakeLists.txt:67-77: src/gaussians_hfsolver.f90 or
! src/gaussians_libint.f90); this file is a third choice with the same module
! name and the same public entry point, so `gaussints` (src/basis.f90:567)
! links against it unchanged:
!
!   getints2(nprim, istart, center, power, coef, alpha, int2)
!
! same arguments as src/gaussians_hfsolver.f90:13-34 (per-function primitive
! counts and 1-based starts, centres(3, n), Cartesian powers(3, n), normalised
! coefficients and exponents; int2 = every unique (ij|kl) in the packed order
! of ijkl2intindex, src/scf.f90:184-223). The integrals come from the MI355X
! engine's store mode (hfx_int2 in include/hfx.h) with every primitive pair
! kept (Schwarz threshold 0). Errors stop with exit code 1 as stop_error does
! (src/utils.f90:171-186).
module gaussians
use iso_c_binding, only: c_int, c_double, c_long_long, c_ptr, c_char, c_associated, c_f_pointer
implicit none
private
public getints2

integer, parameter :: dp = c_double

type, bind(c) :: opts_t
    real(c_double) :: schwarz_threshold
    real(c_double) :: fp32_threshold
    integer(c_int) :: device
    integer(c_int) :: verbose
    integer(c_int) :: kernel_mode
end type

interface
    integer(c_int) function c_default_opts(o) bind(c, name="hfx_default_opts")
    import :: c_int, opts_t
    type(opts_t), intent(out) :: o
    end function

    integer(c_int) function c_create(nbf, nprim, istart, x, y, z, l, m, n, n2, coef, alpha, &
            nat, atno, xyz, o, h) bind(c, name="hfx_create")
    import :: c_int, c_double, c_ptr, opts_t
    integer(c_int), value, intent(in) :: nbf, n2, nat
    integer(c_int), intent(in) :: nprim(nbf), istart(nbf), l(nbf), m(nbf), n(nbf), atno(nat)
    real(c_double), intent(in) :: x(nbf), y(nbf), z(nbf), coef(n2), alpha(n2), xyz(3, nat)
    type(opts_t), intent(in) :: o
    type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function c_int2(h, int2, size) bind(c, name="hfx_int2")
    import :: c_int, c_double, c_long_long, c_ptr
    type(c_ptr), value, intent(in) :: h
    real(c_double), intent(out) :: int2(*)
    integer(c_long_long), value, intent(in) :: size
    end function

    type(c_ptr) function c_last_error(h) bind(c, name="hfx_last_error")
    import :: c_ptr
    type(c_ptr), value, intent(in) :: h
    end function

    subroutine c_destroy(h) bind(c, name="hfx_destroy")
    import :: c_ptr
    type(c_ptr), value, intent(in) :: h
    end subroutine
end interface

contains

subroutine getints2(nprim, istart, center, power, coef, alpha, int2)
integer, intent(in) :: nprim(:), istart(:)
real(dp), intent(in) :: center(:, :)
integer, intent(in) :: power(:, :)
real(dp), intent(in) :: coef(:), alpha(:)
real(dp), intent(out) :: int2(:)

integer(c_int), allocatable :: st0(:), lp(:), mp(:), np(:), atno(:)
real(c_double), allocatable :: xs(:), ys(:), zs(:), atoms(:, :)
type(opts_t) :: o
type(c_ptr) :: h
integer :: n, i, k, nat, stat

n = size(nprim)
allocate(st0(n), lp(n), mp(n), np(n), xs(n), ys(n), zs(n), atoms(3, n))
st0 = istart - 1
lp = power(1, :)
mp = power(2, :)
np = power(3, :)
xs = center(1, :)
ys = center(2, :)
zs = center(3, :)
! the engine groups functions by atom: the distinct centres stand for the
! nuclei (charges are not needed for the two-electron integrals)
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
stat = c_default_opts(o)
o%schwarz_threshold = 0
stat = c_create(n, nprim, st0, xs, ys, zs, lp, mp, np, size(coef), coef, alpha, nat, atno, &
    atoms(:, 1:nat), o, h)
if (stat /= 0) call fail(h)
stat = c_int2(h, int2, int(size(int2), c_long_long))
if (stat /= 0) call fail(h)
call c_destroy(h)
end subroutine

subroutine fail(h)
type(c_ptr), intent(in) :: h
character(kind=c_char), pointer :: s(:)
character(len=512) :: msg
integer :: i
msg = ""
if (c_associated(c_last_error(h))) then
    call c_f_pointer(c_last_error(h), s, [512])
    do i = 1, 512
        if (s(i) == char(0)) exit
        msg(i:i) = s(i)
    end do
end if
print "(a)", "ERROR STOP: getints2: " // trim(msg)
stop 1
end subroutine

end module
