!  ttime's per-step diagnostics from the device (ttime.F90:29-69,136): the
!  subroutines of the reference that ttime calls, with their argument lists,
!  filled from one fribo_diagnostics pass per step instead of full copies of
!  q2/q3 to the host:
!    vmaxv(vmax), vminv(vmin)     vcheck.F90:4-73
!    divgck(qmax)                 divergence.F90:49-79
!    kinen(ektot)                 utils_proc.F90:474-515
!    getintvort(vir)              utils_proc.F90:582-629
!    cfl_diag(cflm, q3avg)        the values cfl (cfl.F90:4-51) leaves in
!                                 module param, for ttime's CALL cfl
!  The driver sets fribo_diag_h once and fribo_diag_valid = .false. after
!  every step; the first of these calls in a step runs the device pass.

      MODULE fribo_ttime

      USE iso_c_binding
      USE fribo_gpu_mod

      implicit none

      type(c_ptr) :: fribo_diag_h = c_null_ptr
      logical :: fribo_diag_valid = .false.
      real(c_double) :: fribo_diag(9)

      contains

        subroutine fribo_diag_refresh()
          integer(c_int) :: rc
          if (fribo_diag_valid) return
          rc = fribo_diagnostics(fribo_diag_h, fribo_diag)
          if (rc /= FRIBO_OK) then
            write(*,*) ' fribo_diagnostics failed: ', trim(fribo_error_message())
            stop 1
          endif
          fribo_diag_valid = .true.
        end subroutine fribo_diag_refresh

        subroutine vmaxv(vmax)
          real(c_double), intent(out) :: vmax(2)
          call fribo_diag_refresh()
          vmax = fribo_diag(1:2)
        end subroutine vmaxv

        subroutine vminv(vmin)
          real(c_double), intent(out) :: vmin(2)
          call fribo_diag_refresh()
          vmin = fribo_diag(3:4)
        end subroutine vminv

        subroutine divgck(qmax)
          real(c_double), intent(out) :: qmax
          call fribo_diag_refresh()
          qmax = fribo_diag(5)
        end subroutine divgck

        subroutine kinen(ektot)
          real(c_double), intent(out) :: ektot
          call fribo_diag_refresh()
          ektot = fribo_diag(6)
        end subroutine kinen

        subroutine getintvort(vir)
          real(c_double), intent(out) :: vir
          call fribo_diag_refresh()
          vir = fribo_diag(7)
        end subroutine getintvort

        subroutine cfl_diag(cflm, q3avg)
          real(c_double), intent(out) :: cflm, q3avg
          call fribo_diag_refresh()
          cflm = fribo_diag(8)
          q3avg = fribo_diag(9)
        end subroutine cfl_diag

      END MODULE fribo_ttime
