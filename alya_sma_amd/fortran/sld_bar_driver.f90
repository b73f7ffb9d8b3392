!-----------------------------------------------------------------------
! Fortran host of the uniaxial SMA bar (10x2x2 HEX08 over 10x2x2 mm,
! displacement-controlled ramp to 6% strain and back): builds the mesh,
! the boundary conditions of tests/uniaxial_test/sm154_d1C.sld.dat:80-107
! (u_x = 0 on x = 0, u_y = 0 on y = 0, u_z = 0 on z = 0, prescribed u_x on
! x = 10 through a DISCRETE,SMOOTH function) and runs the explicit steps on
! the GPU through mod_sld_dev, in place of Alya's time loop
! (Timste/Begste/Doiter/Endste) around sld_explicit.
!
! usage: sld_bar_driver NSTEPS DT SCHEME OUTFILE [NRUN]
!   NSTEPS sets the load cycle length, NRUN (default NSTEPS) the steps run
! Output (stream): real64 u(3,npoin); real64 sdv(3,8,nelem); real64 fint(3,npoin)
! and one line per 100 steps on stdout: step, time, tip u_x, tip reaction.
!-----------------------------------------------------------------------
program sld_bar_driver
  use iso_c_binding
  use mod_sld_dev
  implicit none
  integer(c_int32_t), parameter :: nx = 10, ny = 2, nz = 2
  real(c_double), parameter :: lx = 10.0_8, ly = 2.0_8, lz = 2.0_8, strain = 0.06_8
  integer :: npoin, nelem, ip, istep, nsteps, nrun, uout, scheme
  real(c_double), allocatable, target :: coord(:,:), props(:), rho(:), table(:,:), u(:,:), f(:,:), sdv(:,:,:)
  real(c_double), allocatable, target :: frx(:,:)
  real(c_double) :: ener(5)
  integer(c_int32_t), allocatable, target :: lnods(:,:), func_of_node(:), npts(:), smooth(:)
  integer(c_int8_t), allocatable, target :: fixmask(:)
  type(sld_mesh_desc) :: mesh
  type(sld_material_desc) :: mat
  type(sld_bc_desc) :: bc
  type(c_ptr) :: dev
  real(c_double), target :: t(1)
  real(c_double) :: dt, t_end, tipu, tipf, tol
  character(len=256) :: arg
  integer(c_int) :: rc

  call get_command_argument(1, arg); read(arg, *) nsteps
  call get_command_argument(2, arg); read(arg, *) dt
  call get_command_argument(3, arg); read(arg, *) scheme
  nrun = nsteps
  if (command_argument_count() >= 5) then
     call get_command_argument(5, arg); read(arg, *) nrun
  end if
  npoin = (nx + 1) * (ny + 1) * (nz + 1)
  nelem = nx * ny * nz
  allocate(coord(3, npoin), lnods(8, nelem), fixmask(npoin), func_of_node(npoin))
  call sld_dev_check(sld_mesh_box(nx, ny, nz, lx, ly, lz, c_loc(coord), c_loc(lnods)), 'sld_mesh_box')

  ! boundary conditions (kfl_fixno_sld) and the load function on the x = lx face
  tol = 1.0e-9_8 * lx
  fixmask = 0
  func_of_node = -1
  do ip = 1, npoin
     if (abs(coord(1, ip)) < tol) fixmask(ip) = ior(fixmask(ip), 1_c_int8_t)
     if (abs(coord(2, ip)) < tol) fixmask(ip) = ior(fixmask(ip), 2_c_int8_t)
     if (abs(coord(3, ip)) < tol) fixmask(ip) = ior(fixmask(ip), 4_c_int8_t)
     if (abs(coord(1, ip) - lx) < tol) then
        fixmask(ip) = ior(fixmask(ip), 1_c_int8_t)
        func_of_node(ip) = 0
     end if
  end do
  ! NITINOL: E, nu, eL, AS 400/500, SA 250/150 MPa (parco_sld(1:18) layout)
  allocate(props(18), rho(1))
  props = 0.0_8
  props(1) = 60000.0_8; props(2) = 0.33_8; props(3) = 0.05_8
  props(4) = 400.0_8; props(5) = 500.0_8; props(10) = 250.0_8; props(11) = 150.0_8
  props(16) = 1.0_8; props(17) = 1.0_8
  rho(1) = 6.45e-9_8
  ! DISCRETE,SMOOTH table: 0 -> strain*lx -> 0 over nsteps steps (rows vx,vy,vz,t)
  t_end = nsteps * dt
  allocate(table(4, 3), npts(1), smooth(1))
  table = 0.0_8
  table(4, 2) = 0.5_8 * t_end
  table(4, 3) = t_end
  table(1, 2) = strain * lx
  npts(1) = 3
  smooth(1) = 1

  mesh = sld_mesh_desc(int(npoin, c_int64_t), int(nelem, c_int64_t), c_loc(lnods), c_loc(coord))
  mat = sld_material_desc(1_c_int32_t, c_loc(props), c_loc(rho), c_null_ptr, c_null_ptr, c_null_ptr)
  bc = sld_bc_desc(c_loc(fixmask), c_loc(func_of_node), 1_c_int32_t, 3_c_int32_t, c_loc(npts), &
                   c_loc(smooth), c_loc(table), c_null_ptr, -1_c_int32_t)
  call sld_dev_check(sld_dev_create(mesh, mat, bc, dev), 'sld_dev_create')
  ! energy balance of the run (sld_energy, mod_sld_energy.f90:71-162)
  call sld_dev_check(sld_dev_set_energy(dev, 1_c_int32_t), 'sld_dev_set_energy')

  allocate(u(3, npoin), f(3, npoin), sdv(3, 8, nelem))
  do istep = 1, nrun, 100
     call sld_dev_check(sld_dev_step(dev, dt, int(scheme, c_int32_t), 1.033_8, &
                                     int(min(100, nrun - istep + 1), c_int64_t)), 'sld_dev_step')
     call sld_dev_check(sld_dev_get(dev, SLD_FIELD_U, c_loc(u), int(3*npoin, c_int64_t)), 'get u')
     call sld_dev_check(sld_dev_get(dev, SLD_FIELD_FINT, c_loc(f), int(3*npoin, c_int64_t)), 'get fint')
     call sld_dev_check(sld_dev_get(dev, SLD_FIELD_TIME, c_loc(t), 1_c_int64_t), 'get time')
     tipu = 0.0_8; tipf = 0.0_8
     do ip = 1, npoin
        if (abs(coord(1, ip) - lx) < tol) then
           tipu = max(tipu, u(1, ip))
           tipf = tipf + f(1, ip)
        end if
     end do
     write(*, '(i6,3es16.8)') min(istep + 99, nrun), t(1), tipu, tipf
  end do
  call sld_dev_check(sld_dev_get(dev, SLD_FIELD_SDV, c_loc(sdv), int(24*nelem, c_int64_t)), 'get sdv')
  allocate(frx(3, npoin))
  call sld_dev_check(sld_dev_get(dev, SLD_FIELD_REACTIONS, c_loc(frx), int(3*npoin, c_int64_t)), &
                     'get reactions')
  call sld_dev_check(sld_dev_energy(dev, ener), 'sld_dev_energy')
  write(*, '(a,4es16.8)') 'energies allie allwk allke etota', ener(1:4)
  call get_command_argument(4, arg)
  open(newunit=uout, file=trim(arg), access='stream', form='unformatted', status='replace')
  write(uout) u
  write(uout) sdv
  write(uout) f
  write(uout) frx
  write(uout) ener
  close(uout)
  rc = sld_dev_destroy(dev)
end program sld_bar_driver
