! vv_dropin3.F90 -- Fortran drop-in for the time-step routines of rot3dmg.f
! (3D rotating annulus), over the C ABI of libvvmi (include/vvmi3.h).
!
! A program that calls the reference routines links this file (and
! libvvmi.so) in place of the reference's own subroutines.  The routine
! names and argument lists are the reference's; the COMMON blocks are the
! reference's, dimensioned with the caller's PARAMETERs VV3_KR, VV3_KZ,
! VV3_KP (defaults 33, 33, 160 as in rot3dmg.f:23).  poissoninit creates
! the device run from COMMON /num/ /val/ /radius/, with the caller's dr, dz,
! dp.  The three Poisson solves are exact (the reference iterates its
! multigrid to an absolute residual of 1e-5, tolcon at rot3dmg.f:1562).
!
! Residency (the default).  The device run keeps the state between calls;
! every field carries two coherence bits, host-newer (put before its next
! use; all fields start so) and device-newer.  A routine puts only its
! host-newer inputs and leaves its outputs on the device; update, the last
! routine of a step, copies back the fields the reference program writes
! out -- A, B, C, th, uel, vel, wel (rot3dmg.f:279-319; its time loop reads
! none, only the flux) -- so a step moves seven fields and the flux instead
! of about ninety copies.  The solves run on the device when their
! arguments are the COMMON fields and the array rhsw/rhsv/rhsu filled, and
! through the host arrays otherwise.  vv3_dropin_sync returns every
! device-newer COMMON field; vv3_dropin_host_dirty declares a field the
! caller wrote between calls; vv3_dropin_residency(0), or
! VV_DROPIN_RESIDENT=0 in the environment at poissoninit, selects full
! copies of every routine's inputs and outputs (the caller's view of COMMON
! is then the reference routines' after every call).
!
! Replaced routines (rot3dmg.f): Avorticity 341, Bvorticity 477,
! Cvorticity 601, temperature 736, heat 882, update 964, rhsw 1026,
! rhsv 1066, rhsu 1111, poissoninit 1153, poisson1 1216, poisson2 1325,
! poisson3 1434.
!
! Built with -DVV3_HOR=1 it is the drop-in of rot3dmghor.f (the stationary
! horizontal annulus, model VV3_MODEL_HOR): the same routine names, with
! heat(trans1,trans2) (rot3dmghor.f:913) in place of heat(trans1).

#ifndef VV3_KR
#define VV3_KR 33
#endif
#ifndef VV3_KZ
#define VV3_KZ 33
#endif
#ifndef VV3_KP
#define VV3_KP 160
#endif
#ifndef VV3_HOR
#define VV3_HOR 0
#endif

module vv_dropin3_c
  use iso_c_binding
  implicit none

  type, bind(C) :: vv3_config_t
    integer(c_int) :: nr, nz, np
    real(c_double) :: dt, r0, rmax, zmax, Ra, Re, Ptl, Gr, dr, dz, dp
    integer(c_int) :: model
  end type

  integer(c_int), parameter :: F_AN = 0, F_A = 1, F_AP = 2, F_BN = 3, F_B = 4, F_BP = 5, &
       F_CN = 6, F_C = 7, F_CP = 8, F_THN = 9, F_TH = 10, F_THP = 11, F_UEL = 12, F_VEL = 13, &
       F_WEL = 14, F_UELP = 15, F_VELP = 16, F_WELP = 17, F_URH = 18, F_VRH = 19, F_WRH = 20

  interface
    integer(c_int) function vv3_create(cfg, device, run) bind(C, name="vv3_create")
      import :: c_int, c_ptr, vv3_config_t
      type(vv3_config_t), intent(in) :: cfg
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: run
    end function
    integer(c_int) function vv3_put(run, field, host, ldi, ldj) bind(C, name="vv3_put")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      integer(c_int), value :: field, ldi, ldj
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function vv3_get(run, field, host, ldi, ldj) bind(C, name="vv3_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      integer(c_int), value :: field, ldi, ldj
      real(c_double), intent(inout) :: host(*)
    end function
    integer(c_int) function vv3_vorticity(run, which) bind(C, name="vv3_vorticity")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
      integer(c_int), value :: which
    end function
    integer(c_int) function vv3_temperature(run) bind(C, name="vv3_temperature")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
    end function
    integer(c_int) function vv3_heat(run, t1) bind(C, name="vv3_heat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      real(c_double), intent(out) :: t1
    end function
    integer(c_int) function vv3_fluxes(run, t1, t2) bind(C, name="vv3_fluxes")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      real(c_double), intent(out) :: t1, t2
    end function
    integer(c_int) function vv3_rhs(run, which) bind(C, name="vv3_rhs")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
      integer(c_int), value :: which
    end function
    integer(c_int) function vv3_poisson(run, which) bind(C, name="vv3_poisson")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
      integer(c_int), value :: which
    end function
    integer(c_int) function vv3_update(run) bind(C, name="vv3_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
    end function
    type(c_ptr) function vv_last_error() bind(C, name="vv_last_error")
      import :: c_ptr
    end function
  end interface

  type(c_ptr), save :: g_run = c_null_ptr
  integer(c_int), parameter :: g_ldi = VV3_KR + 1, g_ldj = VV3_KZ + 1
  ! coherence bits per field id, the residency switch, the arrays
  ! rhsw / rhsv / rhsu last filled on the device, the fields update returns
  logical, save :: g_resident = .true.
  logical, save :: g_hdirty(0:20) = .true.
  logical, save :: g_ddirty(0:20) = .false.
  integer(c_intptr_t), save :: g_rhs_loc(0:20) = 0
  integer(c_int), parameter :: g_mirror(7) = [F_A, F_B, F_C, F_TH, F_UEL, F_VEL, F_WEL]

contains

  subroutine vv3_check(st, where)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (st >= 0) return
    call c_f_pointer(vv_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (*, *) 'libvvmi error in ', where, ' status ', st, ': ', msg(1:n)
    stop 2
  end subroutine

  subroutine put(field, x)
    integer(c_int), intent(in) :: field
    real(c_double), intent(in) :: x(*)
    call vv3_check(vv3_put(g_run, field, x, g_ldi, g_ldj), 'vv3_put')
  end subroutine

  subroutine get(field, x)
    integer(c_int), intent(in) :: field
    real(c_double), intent(inout) :: x(*)
    call vv3_check(vv3_get(g_run, field, x, g_ldi, g_ldj), 'vv3_get')
  end subroutine

end module vv_dropin3_c

! ------------------------------------------------------------ poissoninit
subroutine poissoninit(nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin3_c
  implicit none
  integer :: nr, nz, np, nlev
  real(c_double) :: dr, dz, dp
  real(c_double) :: Ra, Re, Ptl, Gr, r0, rmax, dt, drc, dzc, dpc
  integer :: nrc, nzc, npc
  common /val/ Ra, Re, Ptl, Gr
  common /radius/ r0, rmax
  common /num/ dt, drc, dzc, dpc, nrc, nzc, npc
  type(vv3_config_t) :: cfg
  if (nr + 1 > VV3_KR .or. nz + 1 > VV3_KZ .or. np > VV3_KP) then
    write (*, *) 'vv_dropin3: grid exceeds the compiled extents VV3_KR, VV3_KZ, VV3_KP'
    stop 2
  end if
  cfg%nr = nr
  cfg%nz = nz
  cfg%np = np
  cfg%dt = dt
  cfg%r0 = r0
  cfg%rmax = rmax
  cfg%zmax = dz * dble(nz)
  cfg%Ra = Ra
  cfg%Re = Re
  cfg%Ptl = Ptl
  cfg%Gr = Gr
  cfg%dr = dr
  cfg%dz = dz
  cfg%dp = dp
  cfg%model = VV3_HOR
  call vv3_check(vv3_create(cfg, 0_c_int, g_run), 'poissoninit')
  nlev = 1
  call vv3_dropin_env_residency()
  g_hdirty = .true.
  g_ddirty = .false.
  g_rhs_loc = 0
end subroutine

subroutine vv3_dropin_env_residency()
  use vv_dropin3_c
  implicit none
  character(len=8) :: val
  integer :: st
  call get_environment_variable('VV_DROPIN_RESIDENT', val, status=st)
  if (st == 0) g_resident = (val(1:1) /= '0')
end subroutine

! mode 1: residency (the default); 0: full copies.  Switching to copies
! first returns every device-newer field to COMMON; switching to residency
! takes COMMON as the newer copy of every field.
subroutine vv3_dropin_residency(mode)
  use vv_dropin3_c
  implicit none
  integer, intent(in) :: mode
  if (mode == 0 .and. g_resident .and. c_associated(g_run)) call vv3_dropin_sync()
  if (mode /= 0 .and. .not. g_resident) g_hdirty = .true.
  g_resident = (mode /= 0)
end subroutine

! every device-newer COMMON field back to the host
subroutine vv3_dropin_sync()
  use vv_dropin3_c
  implicit none
  if (.not. g_resident) return
  call vv3_dropin_get_state([F_AN, F_A, F_AP, F_BN, F_B, F_BP, F_CN, F_C, F_CP, F_THN, F_TH, F_THP, &
                             F_UEL, F_VEL, F_WEL, F_UELP, F_VELP, F_WELP], 18)
end subroutine

! the caller has written COMMON field `field` (an F_* id) since the last call
subroutine vv3_dropin_host_dirty(field)
  use vv_dropin3_c
  implicit none
  integer, intent(in) :: field
  g_hdirty(field) = .true.
  g_ddirty(field) = .false.
end subroutine

subroutine vv3_dropin_mark_device(names, n)
  use vv_dropin3_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
  integer :: j
  do j = 1, n
    g_ddirty(names(j)) = .true.
    g_hdirty(names(j)) = .false.
  end do
end subroutine

! COMMON fields to the device: host-newer ones when resident, all otherwise
subroutine vv3_dropin_put_state(names, n)
  use vv_dropin3_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
#include "vv_dropin3_state.inc"
  integer :: j
  do j = 1, n
    if (g_resident .and. .not. g_hdirty(names(j))) cycle
    select case (names(j))
    case (F_AN);   call put(F_AN, An)
    case (F_A);    call put(F_A, A)
    case (F_AP);   call put(F_AP, Ap)
    case (F_BN);   call put(F_BN, Bn)
    case (F_B);    call put(F_B, B)
    case (F_BP);   call put(F_BP, Bp)
    case (F_CN);   call put(F_CN, Cn)
    case (F_C);    call put(F_C, C)
    case (F_CP);   call put(F_CP, Cp)
    case (F_THN);  call put(F_THN, thn)
    case (F_TH);   call put(F_TH, th)
    case (F_THP);  call put(F_THP, thp)
    case (F_UEL);  call put(F_UEL, uel)
    case (F_VEL);  call put(F_VEL, vel)
    case (F_WEL);  call put(F_WEL, wel)
    case (F_UELP); call put(F_UELP, uelp)
    case (F_VELP); call put(F_VELP, velp)
    case (F_WELP); call put(F_WELP, welp)
    end select
    g_hdirty(names(j)) = .false.
  end do
end subroutine

! COMMON fields from the device: device-newer ones when resident, all otherwise
subroutine vv3_dropin_get_state(names, n)
  use vv_dropin3_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
#include "vv_dropin3_state.inc"
  integer :: j
  do j = 1, n
    if (g_resident .and. .not. g_ddirty(names(j))) cycle
    select case (names(j))
    case (F_AN);   call get(F_AN, An)
    case (F_A);    call get(F_A, A)
    case (F_AP);   call get(F_AP, Ap)
    case (F_BN);   call get(F_BN, Bn)
    case (F_B);    call get(F_B, B)
    case (F_BP);   call get(F_BP, Bp)
    case (F_CN);   call get(F_CN, Cn)
    case (F_C);    call get(F_C, C)
    case (F_CP);   call get(F_CP, Cp)
    case (F_THN);  call get(F_THN, thn)
    case (F_TH);   call get(F_TH, th)
    case (F_THP);  call get(F_THP, thp)
    case (F_UEL);  call get(F_UEL, uel)
    case (F_VEL);  call get(F_VEL, vel)
    case (F_WEL);  call get(F_WEL, wel)
    case (F_UELP); call get(F_UELP, uelp)
    case (F_VELP); call get(F_VELP, velp)
    case (F_WELP); call get(F_WELP, welp)
    end select
    g_ddirty(names(j)) = .false.
  end do
end subroutine

! a routine's outputs: left on the device when resident, copied back otherwise
subroutine vv3_dropin_out(names, n)
  use vv_dropin3_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
  if (g_resident) then
    call vv3_dropin_mark_device(names, n)
  else
    call vv3_dropin_get_state(names, n)
  end if
end subroutine

! a right-hand side (not in COMMON): put when the host copy is newer; left
! on the device (its array recorded for the solve) or copied back
subroutine vv3_dropin_rhs(which, f, x)
  use vv_dropin3_c
  implicit none
  integer(c_int), intent(in) :: which, f
  real(c_double) :: x(*)
  if (.not. g_resident .or. g_hdirty(f)) then
    call put(f, x)
    g_hdirty(f) = .false.
  end if
  call vv3_check(vv3_rhs(g_run, which), 'rhs')
  if (g_resident) then
    call vv3_dropin_mark_device([f], 1)
    g_rhs_loc(f) = loc(x)
  else
    call get(f, x)
  end if
end subroutine

! ------------------------------------------------------- explicit routines
subroutine Avorticity
  use vv_dropin3_c
  implicit none
  call vv3_dropin_put_state([F_A, F_B, F_C, F_AP, F_TH, F_UEL, F_VEL, F_WEL, F_AN], 9)
  call vv3_check(vv3_vorticity(g_run, 1_c_int), 'Avorticity')
  call vv3_dropin_out([F_AN], 1)
end subroutine

subroutine Bvorticity
  use vv_dropin3_c
  implicit none
  call vv3_dropin_put_state([F_A, F_B, F_C, F_BP, F_TH, F_UEL, F_VEL, F_WEL, F_BN], 9)
  call vv3_check(vv3_vorticity(g_run, 2_c_int), 'Bvorticity')
  call vv3_dropin_out([F_BN], 1)
end subroutine

subroutine Cvorticity
  use vv_dropin3_c
  implicit none
  call vv3_dropin_put_state([F_A, F_B, F_C, F_CP, F_TH, F_UEL, F_VEL, F_WEL, F_CN], 9)
  call vv3_check(vv3_vorticity(g_run, 3_c_int), 'Cvorticity')
  call vv3_dropin_out([F_CN], 1)
end subroutine

subroutine temperature
  use vv_dropin3_c
  implicit none
  call vv3_dropin_put_state([F_TH, F_THP, F_UEL, F_VEL, F_WEL, F_THN], 6)
  call vv3_check(vv3_temperature(g_run), 'temperature')
  call vv3_dropin_out([F_THN], 1)
end subroutine

#if VV3_HOR
! heat(trans1,trans2) of rot3dmghor.f:913: inner- and outer-cylinder fluxes
subroutine heat(trans1, trans2)
  use vv_dropin3_c
  implicit none
  real(c_double) :: trans1, trans2
  call vv3_dropin_put_state([F_TH], 1)
  call vv3_check(vv3_heat(g_run, trans1), 'heat')
  call vv3_check(vv3_fluxes(g_run, trans1, trans2), 'heat')
end subroutine
#else
subroutine heat(trans1)
  use vv_dropin3_c
  implicit none
  real(c_double) :: trans1
  call vv3_dropin_put_state([F_TH], 1)
  call vv3_check(vv3_heat(g_run, trans1), 'heat')
end subroutine
#endif

subroutine rhsw(wrh)
  use vv_dropin3_c
  implicit none
  real(c_double) :: wrh(0:VV3_KR,0:VV3_KZ,0:VV3_KP)
  call vv3_dropin_put_state([F_AN, F_BN], 2)
  call vv3_dropin_rhs(1_c_int, F_WRH, wrh)
end subroutine

subroutine rhsv(vrh)
  use vv_dropin3_c
  implicit none
  real(c_double) :: vrh(0:VV3_KR,0:VV3_KZ,0:VV3_KP)
  call vv3_dropin_put_state([F_AN, F_CN], 2)
  call vv3_dropin_rhs(2_c_int, F_VRH, vrh)
end subroutine

subroutine rhsu(urh)
  use vv_dropin3_c
  implicit none
  real(c_double) :: urh(0:VV3_KR,0:VV3_KZ,0:VV3_KP)
  call vv3_dropin_put_state([F_BN, F_CN, F_WEL], 3)
  call vv3_dropin_rhs(3_c_int, F_URH, urh)
end subroutine

! ------------------------------------------------------------ the solves
subroutine vv3_dropin_solve(which, fpsi, fpsip, frhs, psi, psip, rhs)
  use vv_dropin3_c
  implicit none
  integer(c_int), intent(in) :: which, fpsi, fpsip, frhs
  real(c_double) :: psi(*), psip(*), rhs(*)
#include "vv_dropin3_state.inc"
  integer(c_intptr_t) :: lpsi, lpsip
  if (which == 1) then
    lpsi = loc(wel)
    lpsip = loc(welp)
  else if (which == 2) then
    lpsi = loc(vel)
    lpsip = loc(velp)
  else
    lpsi = loc(uel)
    lpsip = loc(uelp)
  end if
  if (g_resident .and. loc(psi) == lpsi .and. loc(psip) == lpsip .and. loc(rhs) == g_rhs_loc(frhs) &
      .and. g_ddirty(frhs)) then
    call vv3_dropin_put_state([fpsi, fpsip], 2)
    call vv3_check(vv3_poisson(g_run, which), 'poisson')
    call vv3_dropin_mark_device([fpsi, fpsip], 2)
    return
  end if
  ! through the host arrays: device-newer COMMON fields returned first,
  ! COMMON the newer copy afterwards
  if (g_resident) call vv3_dropin_get_state([fpsi, fpsip], 2)
  call put(fpsi, psi); call put(fpsip, psip); call put(frhs, rhs)
  call vv3_check(vv3_poisson(g_run, which), 'poisson')
  call get(fpsi, psi); call get(fpsip, psip)
  if (g_resident) then
    call vv3_dropin_host_dirty(int(fpsi))
    call vv3_dropin_host_dirty(int(fpsip))
    call vv3_dropin_host_dirty(int(frhs))
  end if
end subroutine

subroutine poisson1(psi, psip, rhs, nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin3_c
  implicit none
  real(c_double) :: psi(*), psip(*), rhs(*), dr, dz, dp
  integer :: nr, nz, np, nlev
  call vv3_dropin_solve(1_c_int, F_WEL, F_WELP, F_WRH, psi, psip, rhs)
end subroutine

subroutine poisson2(psi, psip, rhs, nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin3_c
  implicit none
  real(c_double) :: psi(*), psip(*), rhs(*), dr, dz, dp
  integer :: nr, nz, np, nlev
  call vv3_dropin_solve(2_c_int, F_VEL, F_VELP, F_VRH, psi, psip, rhs)
end subroutine

subroutine poisson3(psi, psip, rhs, nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin3_c
  implicit none
  real(c_double) :: psi(*), psip(*), rhs(*), dr, dz, dp
  integer :: nr, nz, np, nlev
  call vv3_dropin_solve(3_c_int, F_UEL, F_UELP, F_URH, psi, psip, rhs)
end subroutine

! ----------------------------------------------------------------- update
subroutine update
  use vv_dropin3_c
  implicit none
  call vv3_dropin_put_state([F_AN, F_A, F_AP, F_BN, F_B, F_BP, F_CN, F_C, F_CP, F_THN, F_TH, F_THP, &
                             F_UEL, F_VEL, F_WEL], 15)
  call vv3_check(vv3_update(g_run), 'update')
  call vv3_dropin_out([F_AN, F_A, F_AP, F_BN, F_B, F_BP, F_CN, F_C, F_CP, F_THN, F_TH, F_THP], 12)
  if (g_resident) call vv3_dropin_get_state(g_mirror, 7)
end subroutine
