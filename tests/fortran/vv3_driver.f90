! vv3_driver.f90 -- test driver of the 3D Fortran drop-in (vv_dropin3.F90).
!
! Replays the time loop of rot3dmg.f (lines 241-253) through the reference
! routine names, which the drop-in maps onto libvvmi.  Reads from stdin:
! nr nz np nsteps dt Ra Re Ptl Gr, then the path of a binary file with the
! initial temperature th(0:nr+1,0:nz+1,0:np) (raw float64, i fastest) and
! the path of the output file, which receives th, uel, vel, wel (each
! (0:nr+1,0:nz+1,0:np)) and trans1 of every step.
program vv3_driver
  implicit real*8(a-h,o-z)
  parameter(kr=33, kz=33, kp=160)
  real*8 An(0:kr,0:kz,0:kp),A(0:kr,0:kz,0:kp),Ap(0:kr,0:kz,0:kp)
  real*8 Bn(0:kr,0:kz,0:kp),B(0:kr,0:kz,0:kp),Bp(0:kr,0:kz,0:kp)
  real*8 Cn(0:kr,0:kz,0:kp),C(0:kr,0:kz,0:kp),Cp(0:kr,0:kz,0:kp)
  real*8 thn(0:kr,0:kz,0:kp),th(0:kr,0:kz,0:kp),thp(0:kr,0:kz,0:kp)
  real*8 uel(0:kr,0:kz,0:kp),vel(0:kr,0:kz,0:kp),wel(0:kr,0:kz,0:kp)
  real*8 uelp(0:kr,0:kz,0:kp),velp(0:kr,0:kz,0:kp),welp(0:kr,0:kz,0:kp)
  real*8 urh(0:kr,0:kz,0:kp),vrh(0:kr,0:kz,0:kp),wrh(0:kr,0:kz,0:kp)
  real*8, allocatable :: trans(:)
  character(len=512) :: fin, fout
  common/val/Ra,Re,Ptl,Gr
  common/sol1/An,A,Ap,Bn,B,Bp,Cn,C,Cp
  common/sol2/thn,th,thp
  common/sol3/uel,vel,wel,uelp,velp,welp
  common/num/dt,dr,dz,dp,nr,nz,np
  common/radius/r0,rmax

  read (*, *) nr, nz, np, nts, dt, Ra, Re, Ptl, Gr
  read (*, '(a)') fin
  read (*, '(a)') fout
  r0 = 0.5d0
  rmax = 1.0d0
  zmax = 0.5d0
  pi = dacos(-1.d0)
  dr = (rmax - r0)/dble(nr)
  dz = zmax/dble(nz)
  dp = 2.d0*pi/dble(np)
  An = 0; A = 0; Ap = 0; Bn = 0; B = 0; Bp = 0; Cn = 0; C = 0; Cp = 0
  thn = 0; th = 0; thp = 0; uel = 0; vel = 0; wel = 0; uelp = 0; velp = 0; welp = 0
  urh = 0; vrh = 0; wrh = 0
  open (11, file=fin, form='unformatted', access='stream', status='old')
  read (11) (((th(i,j,k), i=0,nr+1), j=0,nz+1), k=0,np)
  close (11)
  call poissoninit(nr, nz, np, dr, dz, dp, nlev)
  thp = th
  allocate (trans(nts))
  do kt = 1, nts
    call Avorticity
    call Bvorticity
    call Cvorticity
    call temperature
    call heat(trans1)
    call rhsw(wrh)
    call poisson1(wel, welp, wrh, nr, nz, np, dr, dz, dp, nlev)
    call rhsv(vrh)
    call poisson2(vel, velp, vrh, nr, nz, np, dr, dz, dp, nlev)
    call rhsu(urh)
    call poisson3(uel, uelp, urh, nr, nz, np, dr, dz, dp, nlev)
    call update
    trans(kt) = trans1
  end do
  open (12, file=fout, form='unformatted', access='stream', status='replace')
  write (12) (((th(i,j,k), i=0,nr+1), j=0,nz+1), k=0,np)
  write (12) (((uel(i,j,k), i=0,nr+1), j=0,nz+1), k=0,np)
  write (12) (((vel(i,j,k), i=0,nr+1), j=0,nz+1), k=0,np)
  write (12) (((wel(i,j,k), i=0,nr+1), j=0,nz+1), k=0,np)
  write (12) trans
  close (12)
end program
