! vv_driver.f90 -- test driver of the Fortran drop-in (vv_dropin.F90).
!
! Replays the time loop of velvor2dperiodicrbconv.f (lines 341-388) through
! the reference routine names, which the drop-in maps onto libvvmi.  Reads
! from stdin: nr np nsteps dt Ra Pr width, then the path of a binary file
! with the initial temperature th(0:nr+1,0:np) (raw little-endian float64,
! i fastest) and the path of the output file, which receives th, C, uel, vel
! and the wall fluxes trans1/trans2 of every step.
program vv_driver
  implicit real*8(a-h,o-z)
  parameter(kr=290, kp=1610)
  real*8 An(0:kr,1,0:kp),A(0:kr,1,0:kp),Ap(0:kr,1,0:kp)
  real*8 Bn(0:kr,1,0:kp),B(0:kr,1,0:kp),Bp(0:kr,1,0:kp)
  real*8 Cn(0:kr,1,0:kp),C(0:kr,1,0:kp),Cp(0:kr,1,0:kp)
  real*8 thn(0:kr,1,0:kp),th(0:kr,1,0:kp),thp(0:kr,1,0:kp)
  real*8 uel(0:kr,1,0:kp),vel(0:kr,1,0:kp),wel(0:kr,1,0:kp)
  real*8 uelp(0:kr,1,0:kp),velp(0:kr,1,0:kp),welp(0:kr,1,0:kp)
  real*8 urh(0:kr,1,0:kp),vrh(0:kr,1,0:kp)
  real*8, allocatable :: trans(:,:)
  character(len=512) :: fin, fout
  common/val/Ra,Re,Ptl,Gr
  common/sol1/An,A,Ap,Bn,B,Bp,Cn,C,Cp
  common/sol2/thn,th,thp
  common/sol3/uel,vel,wel,uelp,velp,welp
  common/num/dt,dr,dz,dp,nr,nz,np
  common/radius/r0,rmax

  read (*, *) nr, np, nts, dt, Ra, Ptl, width
  read (*, '(a)') fin
  read (*, '(a)') fout
  nz = 1
  r0 = 0.5d0
  rmax = 1.0d0
  Re = 0.0d0
  Gr = 0.0d0
  dr = (rmax - r0)/dble(nr)
  dz = 0.0d0
  dp = width/dble(np)
  An = 0; A = 0; Ap = 0; Bn = 0; B = 0; Bp = 0; Cn = 0; C = 0; Cp = 0
  thn = 0; th = 0; thp = 0; uel = 0; vel = 0; wel = 0; uelp = 0; velp = 0; welp = 0
  open (11, file=fin, form='unformatted', access='stream', status='old')
  read (11) ((th(i,1,k), i=0,nr+1), k=0,np)
  close (11)
  call vv_dropin_options(width, 1.0d-10)
  call poissoninit(nr, nz, np, dr, dz, dp, nlev)
  Cp = C
  thp = th
  uelp = uel
  velp = vel
  allocate (trans(2, nts))
  do kt = 1, nts
    if (kt == 1) then
      call Cvorticity1
      call temperature1
    else
      call Cvorticity
      call temperature
    end if
    call heat(trans1, trans2)
    call rhsv(vrh)
    call poisson2(vel, velp, vrh, nr, nz, np, dr, dz, dp, nlev)
    call rhsu(urh)
    call poisson3(uel, uelp, urh, nr, nz, np, dr, dz, dp, nlev)
    call update
    trans(1, kt) = trans1
    trans(2, kt) = trans2
  end do
  open (12, file=fout, form='unformatted', access='stream', status='replace')
  write (12) ((th(i,1,k), i=0,nr+1), k=0,np)
  write (12) ((C(i,1,k), i=0,nr+1), k=0,np)
  write (12) ((uel(i,1,k), i=0,nr+1), k=0,np)
  write (12) ((vel(i,1,k), i=0,nr+1), k=0,np)
  write (12) trans
  close (12)
end program
