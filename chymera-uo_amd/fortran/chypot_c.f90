! ISO_C_BINDING interface to libchypot (include/chypot.h) and the state kept
! by the drop-in SETBDY/BDYGEN/POT3/ZAXPHI of potsolve.F.
!
! The reference keeps its solver state in COMMON blocks (/BDY/, /BDY1/ for the
! boundary tables, pot/boundary.f:22-25; BLKTRI's root tables, pot/pot3.f:37-41)
! and stops the program on errors (pot/boundary.f:895-907).  Here the state is
! one library context and one prepared grid per process (module variables
! below), and every non-zero status from the library stops the program with
! the library's message, so the fail-stop behaviour stays.
!
! Environment:
!   CHYPOT_DEVICE  HIP device index for the context (default 0)
!   CHYPOT_COEF    "full" (default): COEF as the reference writes it, every
!                  J,K;  "k2": only the K=2 plane (the one the hydro loop
!                  reads, radhydro/3dhyd-main.f:708,784), the rest of /COEFS/
!                  is left untouched
module chypot_c
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CHY_OK = 0
  integer(c_int), parameter, public :: CHY_COEF_NONE = 0, CHY_COEF_K2 = 1, CHY_COEF_FULL = 2

  type, bind(C), public :: chy_grid_desc
     integer(c_int) :: jmax, kmax, lmax
     type(c_ptr) :: r, z, rhf, zhf
     real(c_double) :: dtheta, pi, grav
  end type chy_grid_desc

  type, bind(C), public :: chy_diag
     real(c_double) :: tmass, phichk
     integer(c_int) :: klocat
  end type chy_diag

  public :: chy_ctx_create, chy_ctx_destroy, chy_setbdy, chy_grid_destroy
  public :: chy_bdygen_host, chy_pot3_host, chy_zaxphi_host, chy_phac, chy_strerror
  public :: chy_shim_setbdy, chy_shim_bdygen, chy_shim_pot3, chy_shim_zaxphi, chy_shim_phac
  public :: chy_shim_release

  interface
     integer(c_int) function chy_ctx_create(device, out) bind(C, name="chy_ctx_create")
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: out
     end function chy_ctx_create

     integer(c_int) function chy_ctx_destroy(ctx) bind(C, name="chy_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function chy_ctx_destroy

     integer(c_int) function chy_setbdy(ctx, desc, isym, out) bind(C, name="chy_setbdy")
       import :: c_int, c_ptr, chy_grid_desc
       type(c_ptr), value :: ctx
       type(chy_grid_desc), intent(in) :: desc
       integer(c_int), value :: isym
       type(c_ptr), intent(out) :: out
     end function chy_setbdy

     integer(c_int) function chy_grid_destroy(g) bind(C, name="chy_grid_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: g
     end function chy_grid_destroy

     integer(c_int) function chy_bdygen_host(g, maxtrm, isym, redge, rho, phi, diag) &
          bind(C, name="chy_bdygen_host")
       import :: c_int, c_ptr, c_double, chy_diag
       type(c_ptr), value :: g
       integer(c_int), value :: maxtrm, isym
       real(c_double), value :: redge
       real(c_double), intent(in) :: rho(*)
       real(c_double), intent(inout) :: phi(*)
       type(chy_diag), intent(inout) :: diag
     end function chy_bdygen_host

     integer(c_int) function chy_pot3_host(g, npoint, rho, phi, coef, coef_mode, diag) &
          bind(C, name="chy_pot3_host")
       import :: c_int, c_ptr, c_double, chy_diag
       type(c_ptr), value :: g
       integer(c_int), value :: npoint
       real(c_double), intent(in) :: rho(*)
       real(c_double), intent(inout) :: phi(*), coef(*)
       integer(c_int), value :: coef_mode
       type(chy_diag), intent(inout) :: diag
     end function chy_pot3_host

     integer(c_int) function chy_zaxphi_host(g, npoint, phi, diag) bind(C, name="chy_zaxphi_host")
       import :: c_int, c_ptr, c_double, chy_diag
       type(c_ptr), value :: g
       integer(c_int), value :: npoint
       real(c_double), intent(inout) :: phi(*)
       type(chy_diag), intent(inout) :: diag
     end function chy_zaxphi_host

     integer(c_int) function chy_phac(g, phac, n) bind(C, name="chy_phac")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: phac(*)
       integer(c_int), value :: n
     end function chy_phac

     type(c_ptr) function chy_strerror(code) bind(C, name="chy_strerror")
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function chy_strerror

     type(c_ptr) function chy_last_device_error() bind(C, name="chy_last_device_error")
       import :: c_ptr
     end function chy_last_device_error
  end interface

  ! process-wide state of the drop-in routines
  type(c_ptr), save :: ctx = c_null_ptr
  type(c_ptr), save :: grid = c_null_ptr
  logical, save :: initialised = .false.   ! SETBDY(0,...) seen (BDYCHK=1, boundary.f:28-29)
  integer, save :: gj = -1, gk = -1, gl = -1, gisym = 0
  real(c_double), allocatable, save :: gr(:), gz(:), grhf(:), gzhf(:)
  real(c_double), save :: gdth = 0.0d0, gpi = 0.0d0, ggrav = 0.0d0

contains

  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n
    s = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, a, [4096])
    n = 0
    do while (n < 4096)
       if (a(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: s)
    s = transfer(a(1:n), s)
  end function c_string

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc == CHY_OK) return
    write(0, '(a)') ' CHYPOT: '//where//': '//c_string(chy_strerror(rc))
    if (rc == 4) write(0, '(a)') ' CHYPOT: '//c_string(chy_last_device_error())
    write(3, '(a)') ' CHYPOT: '//where//': '//c_string(chy_strerror(rc))
    stop 1
  end subroutine check

  integer function env_int(name, default)
    character(len=*), intent(in) :: name
    integer, intent(in) :: default
    character(len=32) :: v
    integer :: st, ios
    env_int = default
    call get_environment_variable(name, v, status=st)
    if (st /= 0) return
    read(v, *, iostat=ios) env_int
    if (ios /= 0) env_int = default
  end function env_int

  integer(c_int) function coef_mode()
    character(len=8) :: v
    integer :: st
    coef_mode = CHY_COEF_FULL
    call get_environment_variable('CHYPOT_COEF', v, status=st)
    if (st == 0 .and. (v == 'k2' .or. v == 'K2')) coef_mode = CHY_COEF_K2
  end function coef_mode

  ! SETBDY(NCALL,ISYM) (pot/boundary.f:2-82).  NCALL=0 initialises; NCALL>0
  ! re-evaluates the boundary for a moved grid.  Both rebuild the prepared
  ! grid from the current /GRID/ and /blok6/ unless they are unchanged (the
  ! reference calls SETBDY once per thread inside a parallel region,
  ! radhydro/io.f:212-216; the callers are serialised by potsolve.F).
  subroutine chy_shim_setbdy(ncall, isym, jmax, kmax, lmax, r, z, rhf, zhf, dtheta, pi, grav)
    integer, intent(in) :: ncall, isym, jmax, kmax, lmax
    real(c_double), intent(in), target :: r(jmax + 2), z(kmax + 2), rhf(jmax + 2), zhf(kmax + 2)
    real(c_double), intent(in) :: dtheta, pi, grav
    type(chy_grid_desc) :: desc
    integer(c_int) :: rc
    if (.not. c_associated(ctx)) then
       call check(chy_ctx_create(int(env_int('CHYPOT_DEVICE', 0), c_int), ctx), 'SETBDY: context')
    end if
    if (ncall == 0) initialised = .true.
    if (c_associated(grid)) then
       if (same_grid()) return
       call check(chy_grid_destroy(grid), 'SETBDY: release grid')
       grid = c_null_ptr
    end if
    desc%jmax = int(jmax, c_int); desc%kmax = int(kmax, c_int); desc%lmax = int(lmax, c_int)
    desc%r = c_loc(r); desc%z = c_loc(z); desc%rhf = c_loc(rhf); desc%zhf = c_loc(zhf)
    desc%dtheta = dtheta; desc%pi = pi; desc%grav = grav
    rc = chy_setbdy(ctx, desc, int(isym, c_int), grid)
    call check(rc, 'SETBDY')
    gj = jmax; gk = kmax; gl = lmax; gisym = isym
    gr = r; gz = z; grhf = rhf; gzhf = zhf
    gdth = dtheta; gpi = pi; ggrav = grav
  contains
    logical function same_grid()
      same_grid = gj == jmax .and. gk == kmax .and. gl == lmax .and. gisym == isym
      if (.not. same_grid) return
      same_grid = all(gr == r) .and. all(gz == z) .and. all(grhf == rhf) .and. all(gzhf == zhf) &
           .and. gdth == dtheta .and. gpi == pi .and. ggrav == grav
    end function same_grid
  end subroutine chy_shim_setbdy

  subroutine require_setbdy()
    if (.not. initialised .or. .not. c_associated(grid)) then
       ! the reference's message and STOP (pot/boundary.f:311,895-907)
       write(3, '(a)') ' SETBDY HAS NOT BEEN CALLED.  IT MUST BE CALLED AT LEAST ONCE'// &
            ' IN ORDER TO INITIALIZE THE BOUNDARY TABLES.'
       write(0, '(a)') ' CHYPOT: BDYGEN/POT3 called before SETBDY(0,ISYM)'
       stop 1
    end if
  end subroutine require_setbdy

  ! BDYGEN(MAXTRM,ISYM,REDGE) (pot/boundary.f:88-908): boundary values in PHI
  ! from RHO; TMASS of /INSIDE/.
  subroutine chy_shim_bdygen(maxtrm, isym, redge, rho, phi, tmass)
    integer, intent(in) :: maxtrm, isym
    real(c_double), intent(in) :: redge
    real(c_double), intent(in) :: rho(*)
    real(c_double), intent(inout) :: phi(*)
    real(c_double), intent(inout) :: tmass
    type(chy_diag) :: d
    call require_setbdy()
    d%tmass = tmass; d%phichk = 0.0d0; d%klocat = 0
    call check(chy_bdygen_host(grid, int(maxtrm, c_int), int(isym, c_int), redge, rho, phi, d), 'BDYGEN')
    tmass = d%tmass
  end subroutine chy_shim_bdygen

  ! POT3(NPOINT,IPRINT) (pot/pot3.f:15-306) with its ZAXPHI call: PHI, COEF,
  ! PHICHK, KLOCAT.
  subroutine chy_shim_pot3(npoint, rho, phi, coef, phichk, klocat)
    integer, intent(in) :: npoint
    real(c_double), intent(in) :: rho(*)
    real(c_double), intent(inout) :: phi(*), coef(*)
    real(c_double), intent(inout) :: phichk
    integer, intent(inout) :: klocat
    type(chy_diag) :: d
    call require_setbdy()
    d%tmass = 0.0d0; d%phichk = phichk; d%klocat = int(klocat, c_int)
    call check(chy_pot3_host(grid, int(npoint, c_int), rho, phi, coef, coef_mode(), d), 'POT3')
    phichk = d%phichk
    klocat = int(d%klocat)
  end subroutine chy_shim_pot3

  ! ZAXPHI(NPOINT,IPRINT) alone (pot/pot3.f:311-399).
  subroutine chy_shim_zaxphi(npoint, phi, phichk, klocat)
    integer, intent(in) :: npoint
    real(c_double), intent(inout) :: phi(*)
    real(c_double), intent(inout) :: phichk
    integer, intent(inout) :: klocat
    type(chy_diag) :: d
    call require_setbdy()
    d%tmass = 0.0d0; d%phichk = phichk; d%klocat = int(klocat, c_int)
    call check(chy_zaxphi_host(grid, int(npoint, c_int), phi, d), 'ZAXPHI')
    phichk = d%phichk
    klocat = int(d%klocat)
  end subroutine chy_shim_zaxphi

  ! PHAC(K), K=2..KMAX+1, of the last POT3/ZAXPHI (for the IPRINT report).
  subroutine chy_shim_phac(phac, n)
    integer, intent(in) :: n
    real(c_double), intent(out) :: phac(n)
    call require_setbdy()
    call check(chy_phac(grid, phac, int(n, c_int)), 'ZAXPHI: PHAC')
  end subroutine chy_shim_phac

  ! Releases the grid and the context (optional; the process end does too).
  subroutine chy_shim_release()
    integer(c_int) :: rc
    if (c_associated(grid)) rc = chy_grid_destroy(grid)
    if (c_associated(ctx)) rc = chy_ctx_destroy(ctx)
    grid = c_null_ptr
    ctx = c_null_ptr
    initialised = .false.
  end subroutine chy_shim_release

end module chypot_c
