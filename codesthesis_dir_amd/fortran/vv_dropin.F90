! vv_dropin.F90 -- Fortran drop-in for the time-step routines of
! velvor2dperiodicrbconv.f, over the C ABI of libvvmi (include/vvmi.h).
!
! A program that calls the reference routines links this file (and
! libvvmi.so) in place of the reference's own subroutines.  The routine
! names and argument lists are the reference's; the COMMON blocks are the
! reference's, dimensioned with the caller's PARAMETERs VV_KR and VV_KP
! (defaults 290 and 1610 as in velvor2dperiodicrbconv.f:54).  The device run
! is created by poissoninit from COMMON /num/ /val/ /radius/.
!
! Residency (the default).  The device run keeps the state between calls and
! every field carries two coherence bits: host-newer (put before the field's
! next use; all fields start so) and device-newer.  A routine puts only its
! host-newer inputs and leaves its outputs on the device; update, the last
! routine of a step, then copies back the fields the reference's time loop
! reads between steps -- C, th, uel and vel: its running means, CFL check
! and output files (velvor2dperiodicrbconv.f:398-470) -- so a step moves
! four fields and the scalars (fluxes, clamp count) instead of every
! routine's inputs and outputs.  poisson2/3 solve on the device when their
! arguments are the COMMON fields and the array rhsv/rhsu filled, and fall
! back to the host copies otherwise.  Callers that read other fields
! between steps call vv_dropin_sync (every device-newer field to COMMON);
! callers that write a field the step uses between calls declare it with
! vv_dropin_host_dirty.  vv_dropin_residency(0), or VV_DROPIN_RESIDENT=0 in
! the environment at poissoninit, selects full copies: each routine then
! copies its inputs from COMMON and its outputs back, and the caller's view
! of COMMON is the reference routines' after every call.
!
! Replaced routines (velvor2dperiodicrbconv.f):
!   Cvorticity1 584, Cvorticity 786, temperature1 1056, temperature 1220,
!   heat 1452, update 1557, rhsv 1629, rhsu 1679, poissoninit 1724,
!   poisson2 1796, poisson3 1930.
!
! Knobs the reference hard-codes are set through vv_dropin_options: the
! periodic width (xx, line 132) and the Poisson tolerance.
!
! Built with -DVV_GEOM=1 (and the extents -DVV_KR=420 -DVV_KP=3600 of
! velvor2dnewsch3.4.f:47) the same file is the drop-in for the rotating-
! annulus plane velvor2dnewsch3.4.f: its routines have the same names and
! arguments (Cvorticity1 596, Cvorticity 797, temperature1 1068, ...); the
! run then takes the annulus geometry with Re from COMMON /val/ and the
! program's Robert coefficient 0.05 (velvor2dnewsch3.4.f:1055, 1454).

#ifndef VV_KR
#define VV_KR 290
#endif
#ifndef VV_KP
#define VV_KP 1610
#endif
#ifndef VV_GEOM
#define VV_GEOM 0
#endif

module vv_dropin_c
  use iso_c_binding
  implicit none

  type, bind(C) :: vv_config_t
    integer(c_int) :: nr, np
    real(c_double) :: dt, r0, rmax, width, Ra, Pr, robert, rtol
    integer(c_int) :: max_cycles, nu1, nu2, stats, geom
    real(c_double) :: Re
    integer(c_int) :: direct_cols
    integer(c_int) :: precision
    integer(c_int) :: lines
  end type

  integer(c_int), parameter :: VV_C = 0, VV_CP = 1, VV_CN = 2, VV_TH = 3, VV_THP = 4, &
       VV_THN = 5, VV_UEL = 6, VV_VEL = 7, VV_UELP = 8, VV_VELP = 9, VV_VRH = 10, VV_URH = 11

  interface
    integer(c_int) function vv_create(cfg, device, run, nlev) bind(C, name="vv_create")
      import :: c_int, c_ptr, vv_config_t
      type(vv_config_t), intent(in) :: cfg
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: run
      integer(c_int), intent(out) :: nlev
    end function
    integer(c_int) function vv_put(run, field, host, ld) bind(C, name="vv_put")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      integer(c_int), value :: field, ld
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function vv_get(run, field, host, ld) bind(C, name="vv_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      integer(c_int), value :: field, ld
      real(c_double), intent(inout) :: host(*)
    end function
    integer(c_int) function vv_cvorticity(run, first) bind(C, name="vv_cvorticity")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
      integer(c_int), value :: first
    end function
    integer(c_int) function vv_temperature(run, first) bind(C, name="vv_temperature")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
      integer(c_int), value :: first
    end function
    integer(c_int) function vv_heat(run, t1, t2) bind(C, name="vv_heat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      real(c_double), intent(out) :: t1, t2
    end function
    integer(c_int) function vv_rhs(run) bind(C, name="vv_rhs")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
    end function
    integer(c_int) function vv_update(run, nclamp) bind(C, name="vv_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: run
      integer(c_int), intent(out) :: nclamp
    end function
    integer(c_int) function vv_poisson(run, which, cycles, relres) bind(C, name="vv_poisson")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      integer(c_int), value :: which
      integer(c_int), intent(out) :: cycles(2)
      real(c_double), intent(out) :: relres(2)
    end function
    integer(c_int) function vv_poisson_host(run, which, psi, psip, rhs, ld, cycles, relres) &
         bind(C, name="vv_poisson_host")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: run
      integer(c_int), value :: which, ld
      real(c_double), intent(inout) :: psi(*), psip(*)
      real(c_double), intent(in) :: rhs(*)
      integer(c_int), intent(out) :: cycles(2)
      real(c_double), intent(out) :: relres(2)
    end function
    type(c_ptr) function vv_last_error() bind(C, name="vv_last_error")
      import :: c_ptr
    end function
  end interface

  type(c_ptr), save :: g_run = c_null_ptr
  real(c_double), save :: g_width = 3.0d0        ! velvor2dperiodicrbconv.f:132
  real(c_double), save :: g_rtol = 1.0d-10
  integer(c_int), parameter :: g_ld = VV_KR + 1
  ! coherence bits per field id (VV_C .. VV_URH), the residency switch, the
  ! arrays rhsv / rhsu last filled on the device, and the fields update
  ! returns to COMMON
  logical, save :: g_resident = .true.
  logical, save :: g_hdirty(0:11) = .true.
  logical, save :: g_ddirty(0:11) = .false.
  integer(c_intptr_t), save :: g_rhs_loc(0:11) = 0
  integer(c_int), parameter :: g_mirror(4) = [VV_C, VV_TH, VV_UEL, VV_VEL]

contains

  subroutine vv_check(st, where)
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

end module vv_dropin_c

subroutine vv_dropin_options(width, rtol)
  use vv_dropin_c
  implicit none
  real(c_double), intent(in) :: width, rtol
  g_width = width
  g_rtol = rtol
end subroutine

! ------------------------------------------------------------ poissoninit
subroutine poissoninit(nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin_c
  implicit none
  integer :: nr, nz, np, nlev
  real(c_double) :: dr, dz, dp
  real(c_double) :: Ra, Re, Ptl, Gr, r0, rmax, dt, drc, dzc, dpc
  integer :: nrc, nzc, npc
  common /val/ Ra, Re, Ptl, Gr
  common /radius/ r0, rmax
  common /num/ dt, drc, dzc, dpc, nrc, nzc, npc
  type(vv_config_t) :: cfg
  integer(c_int) :: nl
  cfg%nr = nr
  cfg%np = np
  cfg%dt = dt
  cfg%r0 = r0
  cfg%rmax = rmax
  cfg%width = g_width
  cfg%Ra = Ra
  cfg%Pr = Ptl
  cfg%rtol = g_rtol
  cfg%max_cycles = 100
  cfg%nu1 = 2
  cfg%nu2 = 2
  cfg%stats = 0
#if VV_GEOM == 1
  cfg%robert = 0.05d0
  cfg%geom = 1
  cfg%Re = Re
#else
  cfg%robert = 0.02d0
  cfg%geom = 0
  cfg%Re = 0.0d0
#endif
  cfg%direct_cols = 0
  cfg%precision = 0
  cfg%lines = 0
  call vv_check(vv_create(cfg, 0_c_int, g_run, nl), 'poissoninit')
  nlev = nl
  call vv_dropin_env_residency()
  g_hdirty = .true.
  g_ddirty = .false.
  g_rhs_loc = 0
end subroutine

subroutine vv_dropin_env_residency()
  use vv_dropin_c
  implicit none
  character(len=8) :: val
  integer :: st
  call get_environment_variable('VV_DROPIN_RESIDENT', val, status=st)
  if (st == 0) g_resident = (val(1:1) /= '0')
end subroutine

! mode 1: residency (the default); 0: full copies.  Switching to copies
! first returns every device-newer field to COMMON; switching to residency
! takes COMMON as the newer copy of every field.
subroutine vv_dropin_residency(mode)
  use vv_dropin_c
  implicit none
  integer, intent(in) :: mode
  if (mode == 0 .and. g_resident .and. c_associated(g_run)) call vv_dropin_sync()
  if (mode /= 0 .and. .not. g_resident) g_hdirty = .true.
  g_resident = (mode /= 0)
end subroutine

! every device-newer COMMON field back to the host
subroutine vv_dropin_sync()
  use vv_dropin_c
  implicit none
  if (.not. g_resident) return
  call vv_dropin_get_state([VV_C, VV_CP, VV_CN, VV_TH, VV_THP, VV_THN, VV_UEL, VV_VEL, VV_UELP, VV_VELP], 10)
end subroutine

! the caller has written COMMON field `field` (a VV_* id) since the last call
subroutine vv_dropin_host_dirty(field)
  use vv_dropin_c
  implicit none
  integer, intent(in) :: field
  g_hdirty(field) = .true.
  g_ddirty(field) = .false.
end subroutine

subroutine vv_dropin_mark_device(names, n)
  use vv_dropin_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
  integer :: j
  do j = 1, n
    g_ddirty(names(j)) = .true.
    g_hdirty(names(j)) = .false.
  end do
end subroutine

! ---------------------------------------------------------- explicit step
subroutine vv_dropin_put_state(names, n)
  use vv_dropin_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
  real(c_double) :: An(0:VV_KR,1,0:VV_KP), A(0:VV_KR,1,0:VV_KP), Ap(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Bn(0:VV_KR,1,0:VV_KP), B(0:VV_KR,1,0:VV_KP), Bp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Cn(0:VV_KR,1,0:VV_KP), C(0:VV_KR,1,0:VV_KP), Cp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: thn(0:VV_KR,1,0:VV_KP), th(0:VV_KR,1,0:VV_KP), thp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: uel(0:VV_KR,1,0:VV_KP), vel(0:VV_KR,1,0:VV_KP), wel(0:VV_KR,1,0:VV_KP)
  real(c_double) :: uelp(0:VV_KR,1,0:VV_KP), velp(0:VV_KR,1,0:VV_KP), welp(0:VV_KR,1,0:VV_KP)
  common /sol1/ An, A, Ap, Bn, B, Bp, Cn, C, Cp
  common /sol2/ thn, th, thp
  common /sol3/ uel, vel, wel, uelp, velp, welp
  integer :: j
  do j = 1, n
    if (g_resident .and. .not. g_hdirty(names(j))) cycle
    select case (names(j))
    case (VV_C);    call vv_check(vv_put(g_run, VV_C, C, g_ld), 'put C')
    case (VV_CP);   call vv_check(vv_put(g_run, VV_CP, Cp, g_ld), 'put Cp')
    case (VV_CN);   call vv_check(vv_put(g_run, VV_CN, Cn, g_ld), 'put Cn')
    case (VV_TH);   call vv_check(vv_put(g_run, VV_TH, th, g_ld), 'put th')
    case (VV_THP);  call vv_check(vv_put(g_run, VV_THP, thp, g_ld), 'put thp')
    case (VV_THN);  call vv_check(vv_put(g_run, VV_THN, thn, g_ld), 'put thn')
    case (VV_UEL);  call vv_check(vv_put(g_run, VV_UEL, uel, g_ld), 'put uel')
    case (VV_VEL);  call vv_check(vv_put(g_run, VV_VEL, vel, g_ld), 'put vel')
    case (VV_UELP); call vv_check(vv_put(g_run, VV_UELP, uelp, g_ld), 'put uelp')
    case (VV_VELP); call vv_check(vv_put(g_run, VV_VELP, velp, g_ld), 'put velp')
    end select
    g_hdirty(names(j)) = .false.
  end do
end subroutine

subroutine vv_dropin_get_state(names, n)
  use vv_dropin_c
  implicit none
  integer, intent(in) :: n
  integer(c_int), intent(in) :: names(n)
  real(c_double) :: An(0:VV_KR,1,0:VV_KP), A(0:VV_KR,1,0:VV_KP), Ap(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Bn(0:VV_KR,1,0:VV_KP), B(0:VV_KR,1,0:VV_KP), Bp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Cn(0:VV_KR,1,0:VV_KP), C(0:VV_KR,1,0:VV_KP), Cp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: thn(0:VV_KR,1,0:VV_KP), th(0:VV_KR,1,0:VV_KP), thp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: uel(0:VV_KR,1,0:VV_KP), vel(0:VV_KR,1,0:VV_KP), wel(0:VV_KR,1,0:VV_KP)
  real(c_double) :: uelp(0:VV_KR,1,0:VV_KP), velp(0:VV_KR,1,0:VV_KP), welp(0:VV_KR,1,0:VV_KP)
  common /sol1/ An, A, Ap, Bn, B, Bp, Cn, C, Cp
  common /sol2/ thn, th, thp
  common /sol3/ uel, vel, wel, uelp, velp, welp
  integer :: j
  do j = 1, n
    if (g_resident .and. .not. g_ddirty(names(j))) cycle
    select case (names(j))
    case (VV_C);    call vv_check(vv_get(g_run, VV_C, C, g_ld), 'get C')
    case (VV_CP);   call vv_check(vv_get(g_run, VV_CP, Cp, g_ld), 'get Cp')
    case (VV_CN);   call vv_check(vv_get(g_run, VV_CN, Cn, g_ld), 'get Cn')
    case (VV_TH);   call vv_check(vv_get(g_run, VV_TH, th, g_ld), 'get th')
    case (VV_THP);  call vv_check(vv_get(g_run, VV_THP, thp, g_ld), 'get thp')
    case (VV_THN);  call vv_check(vv_get(g_run, VV_THN, thn, g_ld), 'get thn')
    case (VV_UEL);  call vv_check(vv_get(g_run, VV_UEL, uel, g_ld), 'get uel')
    case (VV_VEL);  call vv_check(vv_get(g_run, VV_VEL, vel, g_ld), 'get vel')
    case (VV_UELP); call vv_check(vv_get(g_run, VV_UELP, uelp, g_ld), 'get uelp')
    case (VV_VELP); call vv_check(vv_get(g_run, VV_VELP, velp, g_ld), 'get velp')
    end select
    g_ddirty(names(j)) = .false.
  end do
end subroutine

! Cn is returned on the interior rows 1..nr-1 (the rows the reference
! writes); the wall rows keep the caller's values as in the reference.
subroutine vv_dropin_cvort(first)
  use vv_dropin_c
  implicit none
  integer(c_int), intent(in) :: first
  real(c_double) :: An(0:VV_KR,1,0:VV_KP), A(0:VV_KR,1,0:VV_KP), Ap(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Bn(0:VV_KR,1,0:VV_KP), B(0:VV_KR,1,0:VV_KP), Bp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Cn(0:VV_KR,1,0:VV_KP), C(0:VV_KR,1,0:VV_KP), Cp(0:VV_KR,1,0:VV_KP)
  common /sol1/ An, A, Ap, Bn, B, Bp, Cn, C, Cp
  real(c_double) :: dt, dr, dz, dp
  integer :: nr, nz, np
  common /num/ dt, dr, dz, dp, nr, nz, np
  real(c_double), allocatable :: w(:,:,:)
  call vv_dropin_put_state([VV_C, VV_CP, VV_TH, VV_UEL, VV_VEL], 5)
  call vv_check(vv_cvorticity(g_run, first), 'Cvorticity')
  if (g_resident) then
    call vv_dropin_mark_device([VV_CN, VV_C], 2)
    return
  end if
  allocate (w(0:VV_KR,1,0:VV_KP))
  call vv_check(vv_get(g_run, VV_CN, w, g_ld), 'get Cn')
  Cn(1:nr-1,1,0:np) = w(1:nr-1,1,0:np)
  deallocate (w)
  call vv_dropin_get_state([VV_C], 1)
end subroutine

subroutine Cvorticity1
  call vv_dropin_cvort(1)
end subroutine

subroutine Cvorticity
  call vv_dropin_cvort(0)
end subroutine

subroutine vv_dropin_temp(first)
  use vv_dropin_c
  implicit none
  integer(c_int), intent(in) :: first
  real(c_double) :: thn(0:VV_KR,1,0:VV_KP), th(0:VV_KR,1,0:VV_KP), thp(0:VV_KR,1,0:VV_KP)
  common /sol2/ thn, th, thp
  real(c_double) :: dt, dr, dz, dp
  integer :: nr, nz, np
  common /num/ dt, dr, dz, dp, nr, nz, np
  real(c_double), allocatable :: w(:,:,:)
  call vv_dropin_put_state([VV_TH, VV_THP, VV_UEL, VV_VEL], 4)
  call vv_check(vv_temperature(g_run, first), 'temperature')
  if (g_resident) then
    call vv_dropin_mark_device([VV_THN, VV_TH], 2)
    return
  end if
  allocate (w(0:VV_KR,1,0:VV_KP))
  call vv_check(vv_get(g_run, VV_THN, w, g_ld), 'get thn')
  thn(1:nr,1,0:np) = w(1:nr,1,0:np)
  deallocate (w)
  call vv_dropin_get_state([VV_TH], 1)
end subroutine

subroutine temperature1
  call vv_dropin_temp(1)
end subroutine

subroutine temperature
  call vv_dropin_temp(0)
end subroutine

subroutine heat(trans1, trans2)
  use vv_dropin_c
  implicit none
  real(c_double) :: trans1, trans2
  call vv_dropin_put_state([VV_TH], 1)
  call vv_check(vv_heat(g_run, trans1, trans2), 'heat')
end subroutine

subroutine vv_dropin_rhs(field, out)
  use vv_dropin_c
  implicit none
  integer(c_int), intent(in) :: field
  real(c_double) :: out(0:VV_KR,1,0:VV_KP)
  real(c_double) :: dt, dr, dz, dp
  integer :: nr, nz, np
  common /num/ dt, dr, dz, dp, nr, nz, np
  real(c_double), allocatable :: w(:,:,:)
  call vv_dropin_put_state([VV_CN], 1)
  call vv_check(vv_rhs(g_run), 'rhs')
  if (g_resident) then
    ! vrh / urh stay on the device; poisson2 / poisson3 find them there
    ! when they are passed this array
    call vv_dropin_mark_device([field], 1)
    g_rhs_loc(field) = loc(out)
    return
  end if
  allocate (w(0:VV_KR,1,0:VV_KP))
  call vv_check(vv_get(g_run, field, w, g_ld), 'get rhs')
  if (field == VV_VRH) then
    out(1:nr,1,0:np) = w(1:nr,1,0:np)
  else
    out(1:nr-1,1,0:np) = w(1:nr-1,1,0:np)
  end if
  deallocate (w)
end subroutine

subroutine rhsv(vrh)
  use vv_dropin_c
  implicit none
  real(c_double) :: vrh(0:VV_KR,1,0:VV_KP)
  call vv_dropin_rhs(VV_VRH, vrh)
end subroutine

subroutine rhsu(urh)
  use vv_dropin_c
  implicit none
  real(c_double) :: urh(0:VV_KR,1,0:VV_KP)
  call vv_dropin_rhs(VV_URH, urh)
end subroutine

! poisson2 / poisson3: on the device-resident fields when psi, psip are the
! COMMON fields and rhs the array rhsv / rhsu last filled, else through the
! host arrays (device-newer COMMON fields returned first, COMMON the newer
! copy afterwards)
subroutine vv_dropin_solve(which, fpsi, fpsip, frhs, psi, psip, rhs, cyc, rel, st)
  use vv_dropin_c
  implicit none
  integer(c_int), intent(in) :: which, fpsi, fpsip, frhs
  real(c_double) :: psi(0:VV_KR,1,0:VV_KP), psip(0:VV_KR,1,0:VV_KP), rhs(0:VV_KR,1,0:VV_KP)
  integer(c_int), intent(out) :: cyc(2), st
  real(c_double), intent(out) :: rel(2)
  real(c_double) :: uel(0:VV_KR,1,0:VV_KP), vel(0:VV_KR,1,0:VV_KP), wel(0:VV_KR,1,0:VV_KP)
  real(c_double) :: uelp(0:VV_KR,1,0:VV_KP), velp(0:VV_KR,1,0:VV_KP), welp(0:VV_KR,1,0:VV_KP)
  common /sol3/ uel, vel, wel, uelp, velp, welp
  integer(c_intptr_t) :: lpsi, lpsip
  if (which == 2) then
    lpsi = loc(vel)
    lpsip = loc(velp)
  else
    lpsi = loc(uel)
    lpsip = loc(uelp)
  end if
  if (g_resident .and. loc(psi) == lpsi .and. loc(psip) == lpsip .and. loc(rhs) == g_rhs_loc(frhs) &
      .and. g_ddirty(frhs)) then
    call vv_dropin_put_state([fpsi, fpsip], 2)
    st = vv_poisson(g_run, which, cyc, rel)
    if (st >= 0) call vv_dropin_mark_device([fpsi, fpsip], 2)
    return
  end if
  if (g_resident) call vv_dropin_get_state([fpsi, fpsip], 2)
  st = vv_poisson_host(g_run, which, psi, psip, rhs, g_ld, cyc, rel)
  if (g_resident) then
    call vv_dropin_host_dirty(int(fpsi))
    call vv_dropin_host_dirty(int(fpsip))
    call vv_dropin_host_dirty(int(frhs))
  end if
end subroutine

subroutine poisson2(psi, psip, rhs, nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin_c
  implicit none
  integer :: nr, nz, np, nlev
  real(c_double) :: dr, dz, dp
  real(c_double) :: psi(0:VV_KR,1,0:VV_KP), psip(0:VV_KR,1,0:VV_KP), rhs(0:VV_KR,1,0:VV_KP)
  integer(c_int) :: cyc(2), st
  real(c_double) :: rel(2)
  call vv_dropin_solve(2_c_int, VV_VEL, VV_VELP, VV_VRH, psi, psip, rhs, cyc, rel, st)
  call vv_check(st, 'poisson2')
  if (st == 1) then          ! the reference stops after 500 sweeps (line 1843)
    write (*, *) 'poisson2: not converged, relative residual ', rel(1)
    stop 1
  end if
end subroutine

subroutine poisson3(psi, psip, rhs, nr, nz, np, dr, dz, dp, nlev)
  use vv_dropin_c
  implicit none
  integer :: nr, nz, np, nlev
  real(c_double) :: dr, dz, dp
  real(c_double) :: psi(0:VV_KR,1,0:VV_KP), psip(0:VV_KR,1,0:VV_KP), rhs(0:VV_KR,1,0:VV_KP)
  integer(c_int) :: cyc(2), st
  real(c_double) :: rel(2)
  call vv_dropin_solve(3_c_int, VV_UEL, VV_UELP, VV_URH, psi, psip, rhs, cyc, rel, st)
  call vv_check(st, 'poisson3')
  if (st == 1) then
    write (*, *) 'poisson3: not converged, relative residual ', rel(1)
    stop 1
  end if
end subroutine

subroutine update
  use vv_dropin_c
  implicit none
  real(c_double) :: An(0:VV_KR,1,0:VV_KP), A(0:VV_KR,1,0:VV_KP), Ap(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Bn(0:VV_KR,1,0:VV_KP), B(0:VV_KR,1,0:VV_KP), Bp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: Cn(0:VV_KR,1,0:VV_KP), C(0:VV_KR,1,0:VV_KP), Cp(0:VV_KR,1,0:VV_KP)
  real(c_double) :: thn(0:VV_KR,1,0:VV_KP), th(0:VV_KR,1,0:VV_KP), thp(0:VV_KR,1,0:VV_KP)
  common /sol1/ An, A, Ap, Bn, B, Bp, Cn, C, Cp
  common /sol2/ thn, th, thp
  real(c_double) :: dt, dr, dz, dp
  integer :: nr, nz, np
  common /num/ dt, dr, dz, dp, nr, nz, np
  integer(c_int) :: ncl
  call vv_dropin_put_state([VV_C, VV_CN, VV_TH, VV_THN, VV_VEL, VV_UEL], 6)
  call vv_check(vv_update(g_run, ncl), 'update')
  if (ncl > 0) write (*, *) 'WARNING! Temp damped'
  if (g_resident) then
    call vv_dropin_mark_device([VV_C, VV_CP, VV_CN, VV_TH, VV_THP, VV_THN], 6)
    call vv_dropin_get_state(g_mirror, 4)
    return
  end if
  call vv_dropin_get_state([VV_C, VV_CP, VV_TH, VV_THP], 4)
  ! the reference leaves the new level in Cn/thn as well (walls, clamp)
  Cn(0,1,0:np) = C(0,1,0:np)
  Cn(nr,1,0:np) = C(nr,1,0:np)
  thn(0:nr+1,1,0:np) = th(0:nr+1,1,0:np)
end subroutine
