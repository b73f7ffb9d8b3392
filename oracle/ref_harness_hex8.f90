!-----------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY.  Harness program that runs the reference's own
! HEX08 shape functions (mod_lagrange shape3), 2x2x2 Gauss rule
! (mod_quadrature_gauss gauss_qua), Jacobian (jacobi) and 3x3 inverse
! (invmtx) on a batch of elements, all compiled from the reference sources
! by oracle/Makefile.  The loops that combine them restate
! elmca2.f90:105-111 (gpcar, gpvol = w*detJ) and sld_elmpre.f90:149-162,
! 219-226 (F = I + sum_a u_a (x) dN_a/dX, detF by invmtx).
!
! Input  (stream):  int64 n; real64 elcod(3,8,n); real64 eldis(3,8,n)
! Output (stream):  real64 posgp(3,8); real64 weigp(8); real64 shapf(8,8);
!                   real64 deriv(3,8,8);
!                   real64 gpcar(3,8,8,n); real64 gpvol(8,n);
!                   real64 F(3,3,8,n); real64 detF(8,n)
!-----------------------------------------------------------------------
program ref_harness_hex8
  use def_kintyp_basic, only : ip, rp
  use mod_lagrange, only : shape3
  use mod_quadrature_gauss, only : gauss_qua
  implicit none
  character(len=4096) :: fin, fout
  integer :: uin, uout
  integer(8) :: n, ie
  integer(ip) :: ierro, ig, ia, i, j
  real(rp) :: posgp(3,8), weigp(8), shapf(8,8), deriv(3,8,8), heslo(6,8)
  real(rp), allocatable :: elcod(:,:,:), eldis(:,:,:), gpcar(:,:,:,:), gpvol(:,:), F(:,:,:,:), detF(:,:)
  real(rp) :: xjacm(3,3), xjaci(3,3), gpdet, Finv(3,3)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read(uin) n
  allocate(elcod(3,8,n), eldis(3,8,n), gpcar(3,8,8,n), gpvol(8,n), F(3,3,8,n), detF(8,n))
  read(uin) elcod
  read(uin) eldis
  close(uin)

  call gauss_qua(3_ip, 8_ip, posgp, weigp, ierro)
  do ig = 1, 8
     call shape3(posgp(:,ig), 8_ip, shapf(:,ig), deriv(:,:,ig), heslo, ierro)
  end do

  do ie = 1, n
     do ig = 1, 8
        call jacobi(3_ip, 8_ip, elcod(:,:,ie), deriv(:,:,ig), xjacm, xjaci, gpcar(:,:,ig,ie), gpdet)
        gpvol(ig,ie) = weigp(ig) * gpdet
        F(:,:,ig,ie) = 0.0_rp
        do i = 1, 3
           do j = 1, 3
              do ia = 1, 8
                 F(i,j,ig,ie) = F(i,j,ig,ie) + eldis(i,ia,ie) * gpcar(j,ia,ig,ie)
              end do
           end do
        end do
        F(1,1,ig,ie) = F(1,1,ig,ie) + 1.0_rp
        F(2,2,ig,ie) = F(2,2,ig,ie) + 1.0_rp
        F(3,3,ig,ie) = F(3,3,ig,ie) + 1.0_rp
        call invmtx(F(:,:,ig,ie), Finv, detF(ig,ie), 3_ip)
     end do
  end do

  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(uout) posgp
  write(uout) weigp
  write(uout) shapf
  write(uout) deriv
  write(uout) gpcar
  write(uout) gpvol
  write(uout) F
  write(uout) detF
  close(uout)
end program ref_harness_hex8
