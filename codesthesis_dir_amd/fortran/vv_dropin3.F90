! vv_dropin3.F90 -- Fortran drop-in for the time-step routines of rot3dmg.f
! (3D rotating annulus), over the C ABI of libvvmi (include/vvmi3.h).
!
! A program that calls the reference routines links this file (and
! libvvmi.so) in place of the reference's own subroutines.  The routine
! names and argument lists are the reference's; the COMMON blocks are the
! reference's, dimensioned with the caller's PARAMETERs VV3_KR, VV3_KZ,
! VV3_KP (defaults 33, 33, 160 as in rot3dmg.f:23).  Each routine copies the
! fields it reads from COMMON (or its arguments) to the device run, calls
! the HIP library and copies what it writes back, so the caller's view of
! COMMON is the same as with the reference routines.  poissoninit creates
! the device run from COMMON /num/ /val/ /radius/, with the caller's dr, dz,
! dp.  The three Poisson solves are exact (the reference iterates its
! multigrid to an absolute residual of 1e-5, tolcon at rot3dmg.f:1562).
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
end subroutine

! ------------------------------------------------------- explicit routines
subroutine Avorticity
  use vv_dropin3_c
  implicit none
#include "vv_dropin3_state.inc"
  call put(F_A, A); call put(F_B, B); call put(F_C, C); call put(F_AP, Ap); call put(F_TH, th)
  call put(F_UEL, uel); call put(F_VEL, vel); call put(F_WEL, wel); call put(F_AN, An)
  call vv3_check(vv3_vorticity(g_run, 1_c_int), 'Avorticity')
  call get(F_AN, An)
end subroutine

subroutine Bvorticity
  use vv_dropin3_c
  implicit none
#include "vv_dropin3_state.inc"
  call put(F_A, A); call put(F_B, B); call put(F_C, C); call put(F_BP, Bp); call put(F_TH, th)
  call put(F_UEL, uel); call put(F_VEL, vel); call put(F_WEL, wel); call put(F_BN, Bn)
  call vv3_check(vv3_vorticity(g_run, 2_c_int), 'Bvorticity')
  call get(F_BN, Bn)
end subroutine

subroutine Cvorticity
  use vv_dropin3_c
  implicit none
#include "vv_dropin3_state.inc"
  call put(F_A, A); call put(F_B, B); call put(F_C, C); call put(F_CP, Cp); call put(F_TH, th)
  call put(F_UEL, uel); call put(F_VEL, vel); call put(F_WEL, wel); call put(F_CN, Cn)
  call vv3_check(vv3_vorticity(g_run, 3_c_int), 'Cvorticity')
  call get(F_CN, Cn)
end subroutine

subroutine temperature
  use vv_dropin3_c
  implicit none
#include "vv_dropin3_state.inc"
  call put(F_TH, th); call put(F_THP, thp); call put(F_UEL, uel); call put(F_VEL, vel)
  call put(F_WEL, wel); call put(F_THN, thn)
  call vv3_check(vv3_temperature(g_run), 'temperature')
  call get(F_THN, thn)
end subroutine

#if VV3_HOR
! heat(trans1,trans2) of rot3dmghor.f:913: inner- and outer-cylinder fluxes
subroutine heat(trans1, trans2)
  use vv_dropin3_c
  implicit none
  real(c_double) :: trans1, trans2
#include "vv_dropin3_state.inc"
  call put(F_TH, th)
  call vv3_check(vv3_heat(g_run, trans1), 'heat')
  call vv3_check(vv3_fluxes(g_run, trans1, trans2), 'heat')
end subroutine
#else
subroutine heat(trans1)
  use vv_dropin3_c
  implicit none
  real(c_double) :: trans1
#include "vv_dropin3_state.inc"
  call put(F_TH, th)
  call vv3_check(vv3_heat(g_run, trans1), 'heat')
end subroutine
#endif

subroutine rhsw(wrh)
  use vv_dropin3_c
  implicit none
  real(c_double) :: wrh(0:VV3_KR,0:VV3_KZ,0:VV3_KP)
#include "vv_dropin3_state.inc"
  call put(F_AN, An); call put(F_BN, Bn); call put(F_WRH, wrh)
  call vv3_check(vv3_rhs(g_run, 1_c_int), 'rhsw')
  call get(F_WRH, wrh)
end subroutine

subroutine rhsv(vrh)
  use vv_dropin3_c
  implicit none
  real(c_double) :: vrh(0:VV3_KR,0:VV3_KZ,0:VV3_KP)
#include "vv_dropin3_state.inc"
  call put(F_AN, An); call put(F_CN, Cn); call put(F_VRH, vrh)
  call vv3_check(vv3_rhs(g_run, 2_c_int), 'rhsv')
  call get(F_VRH, vrh)
end subroutine

subroutine rhsu(urh)
  use vv_dropin3_c
  implicit none
  real(c_double) :: urh(0:VV3_KR,0:VV3_KZ,0:VV3_KP)
#include "vv_dropin3_state.inc"
  call put(F_BN, Bn); call put(F_CN, Cn); call put(F_WEL, wel); call put(F_URH, urh)
  call vv3_check(vv3_rhs(g_run, 3_c_int), 'rhsu')
  call get(F_URH, urh)
end subroutine

! ------------------------------------------------------------ the solves
subroutine vv3_dropin_solve(which, fpsi, fpsip, frhs, psi, psip, rhs)
  use vv_dropin3_c
  implicit none
  integer(c_int), intent(in) :: which, fpsi, fpsip, frhs
  real(c_double) :: psi(*), psip(*), rhs(*)
  call put(fpsi, psi); call put(fpsip, psip); call put(frhs, rhs)
  call vv3_check(vv3_poisson(g_run, which), 'poisson')
  call get(fpsi, psi); call get(fpsip, psip)
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
#include "vv_dropin3_state.inc"
  call put(F_AN, An); call put(F_A, A); call put(F_AP, Ap)
  call put(F_BN, Bn); call put(F_B, B); call put(F_BP, Bp)
  call put(F_CN, Cn); call put(F_C, C); call put(F_CP, Cp)
  call put(F_THN, thn); call put(F_TH, th); call put(F_THP, thp)
  call put(F_UEL, uel); call put(F_VEL, vel); call put(F_WEL, wel)
  call vv3_check(vv3_update(g_run), 'update')
  call get(F_AN, An); call get(F_A, A); call get(F_AP, Ap)
  call get(F_BN, Bn); call get(F_B, B); call get(F_BP, Bp)
  call get(F_CN, Cn); call get(F_C, C); call get(F_CP, Cp)
  call get(F_THN, thn); call get(F_TH, th); call get(F_THP, thp)
end subroutine
