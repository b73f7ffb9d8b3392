!-----------------------------------------------------------------------
! mod_sld_dev: iso_c_binding interfaces to the MI355X explicit SMA step
! (C ABI declared in include/alya_sma.h).  A Fortran host such as Alya's
! solidz module uses these to drive the device in place of
! sld_explicit/sld_explicit_tw (mod_sld_solution_methods.f90:101-223,
! :385-574) and the element loop under them.
!-----------------------------------------------------------------------
module mod_sld_dev
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: SLD_MODE_FP64 = 0, SLD_MODE_FP32 = 1, SLD_MODE_SURROGATE = 2
  integer(c_int32_t), parameter :: SLD_FIELD_U = 0, SLD_FIELD_V = 1, SLD_FIELD_A = 2, &
       SLD_FIELD_FINT = 3, SLD_FIELD_MASS = 4, SLD_FIELD_SDV = 5, SLD_FIELD_STRESS = 6, &
       SLD_FIELD_FLAGS = 7, SLD_FIELD_TIME = 8, SLD_FIELD_COUNTERS = 9, SLD_FIELD_PK1 = 10, &
       SLD_FIELD_REACTIONS = 11, SLD_FIELD_HEALTH = 12
  integer(c_int32_t), parameter :: SLD_SCHEME_CD = 0, SLD_SCHEME_TW = 1

  type, bind(C) :: sld_mesh_desc
     integer(c_int64_t) :: npoin, nelem
     type(c_ptr)        :: lnods, coord
  end type sld_mesh_desc

  type, bind(C) :: sld_material_desc
     integer(c_int32_t) :: nmat
     type(c_ptr)        :: props, rho, mat_of_elem, E_of_elem, law
  end type sld_material_desc

  type, bind(C) :: sld_bc_desc
     type(c_ptr)        :: fixmask, func_of_node
     integer(c_int32_t) :: nfunc, func_maxpts
     type(c_ptr)        :: func_npts, func_smooth, func_table, fext0
     integer(c_int32_t) :: fext_func
  end type sld_bc_desc

  ! Neumann loads on QUA04 boundary faces (sld_bouope.f90): code 3 pressure,
  ! 6 traction; lnodb 0-based node indices
  type, bind(C) :: sld_face_desc
     integer(c_int64_t) :: nface
     type(c_ptr)        :: lnodb, code, value, func
     integer(c_int32_t) :: transient, follower
  end type sld_face_desc

  interface
     function sma_last_error() bind(C, name="sma_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function sma_last_error

     function sma160_eval(n_gp, n_mat, props, mat_of_gp, E_of_gp, F, T, sdv_old, sdv_new, S, &
                          flags, precision) bind(C, name="sma160_eval") result(rc)
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int64_t), value :: n_gp
       integer(c_int32_t), value :: n_mat, precision
       type(c_ptr), value :: props, mat_of_gp, E_of_gp, F, T, sdv_old, sdv_new, S, flags
       integer(c_int) :: rc
     end function sma160_eval

     function sld_dev_create(mesh, mat, bc, dev) bind(C, name="sld_dev_create") result(rc)
       import :: c_int, c_ptr, sld_mesh_desc, sld_material_desc, sld_bc_desc
       type(sld_mesh_desc), intent(in) :: mesh
       type(sld_material_desc), intent(in) :: mat
       type(sld_bc_desc), intent(in) :: bc
       type(c_ptr), intent(out) :: dev
       integer(c_int) :: rc
     end function sld_dev_create

     function sld_dev_set_mode(dev, mode) bind(C, name="sld_dev_set_mode") result(rc)
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: dev
       integer(c_int32_t), value :: mode
       integer(c_int) :: rc
     end function sld_dev_set_mode

     function sld_dev_set_faces(dev, faces) bind(C, name="sld_dev_set_faces") result(rc)
       import :: c_int, c_ptr, sld_face_desc
       type(c_ptr), value :: dev
       type(sld_face_desc), intent(in) :: faces
       integer(c_int) :: rc
     end function sld_dev_set_faces

     ! g(3) = grnor_sld * gravi_sld (sld_reaphy.f90:209-223)
     function sld_dev_set_gravity(dev, g) bind(C, name="sld_dev_set_gravity") result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: dev
       real(c_double), intent(in) :: g(3)
       integer(c_int) :: rc
     end function sld_dev_set_gravity

     ! ab(2, nmat) = dampi_sld(1:2, imate), RAYALPHA and RAYBETA
     ! (sld_reaphy.f90:1046-1049; sld_elmmat.f90:160-194)
     function sld_dev_set_damping(dev, ab) bind(C, name="sld_dev_set_damping") result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: dev
       real(c_double), intent(in) :: ab(*)
       integer(c_int) :: rc
     end function sld_dev_set_damping

     ! func = 0-based load function of bodyf_sld (BODYF, sld_reabcs.f90:523), -1 = none
     function sld_dev_set_body_force(dev, func) bind(C, name="sld_dev_set_body_force") result(rc)
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: dev
       integer(c_int32_t), value :: func
       integer(c_int) :: rc
     end function sld_dev_set_body_force

     function sld_dev_set_energy(dev, on) bind(C, name="sld_dev_set_energy") result(rc)
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: dev
       integer(c_int32_t), value :: on
       integer(c_int) :: rc
     end function sld_dev_set_energy

     ! out(5) = allie, allwk, allke, etota, steps (mod_sld_energy.f90:141-162)
     function sld_dev_energy(dev, out) bind(C, name="sld_dev_energy") result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: dev
       real(c_double), intent(out) :: out(5)
       integer(c_int) :: rc
     end function sld_dev_energy

     function sld_dev_step(dev, dt, scheme, phi, nsteps) bind(C, name="sld_dev_step") result(rc)
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: dev
       real(c_double), value :: dt, phi
       integer(c_int32_t), value :: scheme
       integer(c_int64_t), value :: nsteps
       integer(c_int) :: rc
     end function sld_dev_step

     function sld_dev_get(dev, field, host, count) bind(C, name="sld_dev_get") result(rc)
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: dev, host
       integer(c_int32_t), value :: field
       integer(c_int64_t), value :: count
       integer(c_int) :: rc
     end function sld_dev_get

     function sld_dev_set(dev, field, host, count) bind(C, name="sld_dev_set") result(rc)
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: dev, host
       integer(c_int32_t), value :: field
       integer(c_int64_t), value :: count
       integer(c_int) :: rc
     end function sld_dev_set

     function sld_dev_critical_dt(dev) bind(C, name="sld_dev_critical_dt") result(dt)
       import :: c_double, c_ptr
       type(c_ptr), value :: dev
       real(c_double) :: dt
     end function sld_dev_critical_dt

     function sld_dev_sync(dev) bind(C, name="sld_dev_sync") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: dev
       integer(c_int) :: rc
     end function sld_dev_sync

     function sld_dev_destroy(dev) bind(C, name="sld_dev_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: dev
       integer(c_int) :: rc
     end function sld_dev_destroy

     function sld_mesh_box(nx, ny, nz, lx, ly, lz, coord, lnods) bind(C, name="sld_mesh_box") result(rc)
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: nx, ny, nz
       real(c_double), value :: lx, ly, lz
       type(c_ptr), value :: coord, lnods
       integer(c_int) :: rc
     end function sld_mesh_box
  end interface

contains

  ! Text of the last library error, for messages.
  function sld_dev_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    type(c_ptr) :: p
    p = sma_last_error()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function sld_dev_error

  ! Abort with the library's message on a non-zero status (the reference
  ! aborts through runend in the same situations).
  subroutine sld_dev_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
       write(*, '(a,a,a,i0,a,a)') 'SLD_DEV: ', what, ' failed (status ', rc, '): ', sld_dev_error()
       error stop 1
    end if
  end subroutine sld_dev_check

end module mod_sld_dev
