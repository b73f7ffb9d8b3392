!-----------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY.  Harness program that drives the reference's
! own SMA law (module sld_stress_model_160, compiled from the reference
! sources by oracle/Makefile) on a batch of Gauss points read from a raw
! binary file, and writes the results to another raw binary file.
!
! Input  (stream, little endian):
!   int64  n
!   real64 props(18,n)      material parameters per point
!   real64 F(9,n)           deformation gradient, column-major 3x3
!   real64 sdv_old(3,n)     (xS, F_AS, F_SA) at TIME_N
! Output:
!   real64 sdv_new(3,n)     (xS, F_AS, F_SA) written by the law (ITER_K)
!   real64 S(9,n)           second Piola-Kirchhoff stress, column-major
!
! Before every point the harness prints "#PT <i>" on stdout so that the
! warnings the reference prints (Newton non-convergence, eigen-solver
! non-convergence, inconsistent flags) can be attributed to a point.  With a
! third argument "quiet" it prints no markers and reports the wall time of
! the point loop as "#TIME <seconds>" (speed calibration of the CPU oracle).
!
! The law is called exactly as sld_builtin_materials.f90:327-335 calls it,
! with pgaus=1 and flagt=0 (explicit scheme, sld_elmcla.f90:93-97).
!-----------------------------------------------------------------------
program ref_harness_sm160
  use def_kintyp, only : ip, rp
  use sld_stress_model_160, only : eval_sld_stress_model_160
  implicit none
  character(len=4096) :: fin, fout
  character(len=16) :: mode
  logical :: quiet
  integer :: uin, uout
  integer(8) :: n, i, c0, c1, rate
  real(rp), allocatable :: props(:,:), F(:,:), sdv_old(:,:), sdv_new(:,:), S(:,:)
  real(rp) :: gpgdi(3,3,1), gptmp(1), gpsdv(3,1,2), gpstr(3,3,1), gpdds(3,3,3,3,1)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, mode)
  quiet = trim(mode) == 'quiet'
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read(uin) n
  allocate(props(18,n), F(9,n), sdv_old(3,n), sdv_new(3,n), S(9,n))
  read(uin) props
  read(uin) F
  read(uin) sdv_old
  close(uin)

  gptmp = 0.0_rp
  call system_clock(c0, rate)
  do i = 1, n
     if (.not. quiet) write(*,'(a,i0)') '#PT ', i - 1
     gpgdi(:,:,1) = reshape(F(:,i), (/3,3/))
     gpsdv(:,1,1) = 0.0_rp
     gpsdv(:,1,2) = sdv_old(:,i)
     gpstr = 0.0_rp
     call eval_sld_stress_model_160(1_ip, props(:,i), gpgdi, gptmp, gpsdv, gpstr, gpdds, 0_ip)
     sdv_new(:,i) = gpsdv(:,1,1)
     S(:,i) = reshape(gpstr(:,:,1), (/9/))
  end do
  call system_clock(c1)
  if (quiet) write(*,'(a,es12.5)') '#TIME ', real(c1 - c0, rp) / real(rate, rp)

  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(uout) sdv_new
  write(uout) S
  close(uout)
end program ref_harness_sm160
