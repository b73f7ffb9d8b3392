!-----------------------------------------------------------------------
! Drop-in for the reference's SMA law module on the GPU.
!
! eval_sld_stress_model_160 keeps the exact argument list of the
! reference procedure (sld_stress_model_160.f90:36-48), so its caller
! (sld_builtin_materials.f90:327-335) switches by changing one `use` line:
!     use sld_stress_model_160      ->   use sld_stress_model_160_dev
! The Gauss points of the call are evaluated by the HIP kernel behind
! sma160_eval (include/alya_sma.h).  Only the explicit scheme (flagt = 0)
! is served: the numerical tangent of :118-155 is not on this path.
!-----------------------------------------------------------------------
module sld_stress_model_160_dev
  use iso_c_binding
  use mod_sld_dev
  implicit none
  private
  public :: eval_sld_stress_model_160

contains

  subroutine eval_sld_stress_model_160(pgaus, props, gpgdi, gptmp, gpsdv, gpstr, gpdds, flagt)
    integer,  intent(in)            :: pgaus
    real(8),  intent(in)            :: props(:)
    real(8),  intent(in)            :: gpgdi(:,:,:)
    real(8),  intent(in)            :: gptmp(:)
    real(8),  intent(inout)         :: gpsdv(:,:,:)
    real(8),  intent(out)           :: gpstr(:,:,:)
    real(8),  intent(out)           :: gpdds(:,:,:,:,:)
    integer,  intent(in)            :: flagt
    ! gptmp is accepted and unused (the reference hard-codes gp_tmp = 0, :105);
    ! gpdds is only written by the implicit tangent path, which is not served
    ! per-call staging on the stack (pgaus is an element's Gauss points);
    ! sma160_eval keeps its device and pinned host blocks between calls
    real(c_double), target :: F(max(pgaus, 1), 9), sdv_old(max(pgaus, 1), 3), &
                              sdv_new(max(pgaus, 1), 3), S(max(pgaus, 1), 9), pr(18)
    integer(c_int32_t), target :: flags(max(pgaus, 1))
    integer :: g, i, j
    integer(c_int) :: rc

    if (flagt /= 0) then
       write(*,*) 'eval_sld_stress_model_160 (device): only the explicit scheme (flagt = 0) is supported'
       error stop 1
    end if
    if (size(props) < 18) then
       write(*,*) 'eval_sld_stress_model_160 (device): props needs 18 entries'
       error stop 1
    end if
    if (pgaus <= 0) return
    ! SoA: component k of point g at (g, k), k = i + 3*(j-1) (column-major 3x3)
    pr = props(1:18)
    do g = 1, pgaus
       do j = 1, 3
          do i = 1, 3
             F(g, i + 3*(j - 1)) = gpgdi(i, j, g)
          end do
       end do
       sdv_old(g, 1:3) = gpsdv(1:3, g, 2)
    end do
    rc = sma160_eval(int(pgaus, c_int64_t), 1_c_int32_t, c_loc(pr), c_null_ptr, c_null_ptr, &
                     c_loc(F), c_null_ptr, c_loc(sdv_old), c_loc(sdv_new), c_loc(S), &
                     c_loc(flags), 0_c_int32_t)
    call sld_dev_check(rc, 'sma160_eval')
    do g = 1, pgaus
       gpsdv(1:3, g, 1) = sdv_new(g, 1:3)
       do j = 1, 3
          do i = 1, 3
             gpstr(i, j, g) = S(g, i + 3*(j - 1))
          end do
       end do
    end do
  end subroutine eval_sld_stress_model_160

end module sld_stress_model_160_dev
