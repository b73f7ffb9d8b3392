!-----------------------------------------------------------------------
! Batch driver of the device law through the drop-in module: reads Gauss
! points from a raw binary file and writes the results, calling
! eval_sld_stress_model_160 exactly as sld_builtin_materials.f90:327-335
! does (all points of the file in one call, flagt = 0).
!
! Input  (stream): int64 n; real64 props(18,n) (first column used);
!                  real64 F(9,n) column-major per point; real64 sdv_old(3,n)
! Output (stream): real64 sdv_new(3,n); real64 S(9,n)
!-----------------------------------------------------------------------
program sld_law_driver
  use sld_stress_model_160_dev, only : eval_sld_stress_model_160
  implicit none
  character(len=4096) :: fin, fout
  integer :: uin, uout
  integer(8) :: n, i
  real(8), allocatable :: props(:,:), F(:,:), sdv_old(:,:), gpgdi(:,:,:), gpsdv(:,:,:), gpstr(:,:,:)
  real(8), allocatable :: gptmp(:), gpdds(:,:,:,:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read(uin) n
  allocate(props(18,n), F(9,n), sdv_old(3,n))
  read(uin) props
  read(uin) F
  read(uin) sdv_old
  close(uin)
  allocate(gpgdi(3,3,n), gpsdv(3,n,2), gpstr(3,3,n), gptmp(n), gpdds(3,3,3,3,1))
  do i = 1, n
     gpgdi(:,:,i) = reshape(F(:,i), (/3,3/))
     gpsdv(:,i,2) = sdv_old(:,i)
  end do
  gpsdv(:,:,1) = 0.0_8
  gptmp = 0.0_8
  call eval_sld_stress_model_160(int(n), props(:,1), gpgdi, gptmp, gpsdv, gpstr, gpdds, 0)
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(uout) gpsdv(:,:,1)
  write(uout) reshape(gpstr, (/9*n/))
  close(uout)
end program sld_law_driver
