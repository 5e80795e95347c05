!  iso_c_binding interface to libfribo_gpu (include/fribo_gpu.h).
!
!  The Fortran driver of FRiBoIS2D keeps its module state (param,
!  local_arrays, outflow_vars, rbm, mls_param) in host arrays with the
!  reference extents and calls these entry points in place of the body of
!  fluid_tsch (src/fluid_tsch.F90:25-166).  Arrays are passed by reference
!  in their Fortran layout; logicals travel as 0/1 integers.

      MODULE fribo_gpu_mod

      USE iso_c_binding

      implicit none

      integer(c_int), parameter :: FRIBO_OK = 0
      integer(c_int), parameter :: FRIBO_ERR_DT_TOO_SMALL = 1
      integer(c_int), parameter :: FRIBO_ERR_CFL_TOO_LARGE = 2
      integer(c_int), parameter :: FRIBO_ERR_MARKER_OUTSIDE = 3
      integer(c_int), parameter :: FRIBO_ERR_PARTITION_UNITY = 4
      integer(c_int), parameter :: FRIBO_ERR_ARG = 5
      integer(c_int), parameter :: FRIBO_ERR_HIP = 6
      integer(c_int), parameter :: FRIBO_ERR_LIMIT = 7

      type, bind(c) :: fribo_params
        integer(c_int32_t) :: n2, n3
        integer(c_int32_t) :: nsst, idtv, ibact, forst, wfun, inslws, inslwn
        real(c_double) :: alx3, rext2, visc, cou, q3uw, q3lw
        real(c_double) :: dtmax, cflmax, cfllim, radiusSD
      end type fribo_params

      interface

        integer(c_int) function fribo_init(p, rc, zz, zmm, zmmt, ldz, wr, h) &
            bind(c, name="fribo_init")
          import :: c_int, c_double, c_ptr, fribo_params
          type(fribo_params), intent(in) :: p
          real(c_double), intent(in) :: rc(*), zz(*), zmm(*), zmmt(*), wr(*)
          integer(c_int), value :: ldz
          type(c_ptr), intent(out) :: h
        end function fribo_init

        integer(c_int) function fribo_finalize(h) bind(c, name="fribo_finalize")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function fribo_finalize

        integer(c_int) function fribo_set_state(h, q2, q3, pr, ru2, ru3, qb2s, qb3s, &
            qb2n, qb3n, dq2x2o, dq3x2o) bind(c, name="fribo_set_state")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: h
          real(c_double), intent(in) :: q2(*), q3(*), pr(*), ru2(*), ru3(*)
          real(c_double), intent(in) :: qb2s(*), qb3s(*), qb2n(*), qb3n(*)
          real(c_double), intent(in) :: dq2x2o(*), dq3x2o(*)
        end function fribo_set_state

        integer(c_int) function fribo_get_state(h, q2, q3, pr, ru2, ru3, qb2s, qb3s, &
            qb2n, qb3n, dq2x2o, dq3x2o) bind(c, name="fribo_get_state")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: h
          real(c_double), intent(inout) :: q2(*), q3(*), pr(*), ru2(*), ru3(*)
          real(c_double), intent(inout) :: qb2s(*), qb3s(*), qb2n(*), qb3n(*)
          real(c_double), intent(inout) :: dq2x2o(*), dq3x2o(*)
        end function fribo_get_state

        integer(c_int) function fribo_set_markers(h, nlm, lmcc, lmvel, lmA) &
            bind(c, name="fribo_set_markers")
          import :: c_int, c_int32_t, c_double, c_ptr
          type(c_ptr), value :: h
          integer(c_int32_t), value :: nlm
          real(c_double), intent(in) :: lmcc(*), lmvel(*), lmA(*)
        end function fribo_set_markers

        integer(c_int) function fribo_fluid_tsch(h, ns, dt, dt_o, cflm) &
            bind(c, name="fribo_fluid_tsch")
          import :: c_int, c_int32_t, c_double, c_ptr
          type(c_ptr), value :: h
          integer(c_int32_t), value :: ns
          real(c_double), intent(inout) :: dt
          real(c_double), value :: dt_o
          real(c_double), intent(out) :: cflm
        end function fribo_fluid_tsch

        integer(c_int) function fribo_cfl_dt(h, ns, dt, dt_o, cflm) &
            bind(c, name="fribo_cfl_dt")
          import :: c_int, c_int32_t, c_double, c_ptr
          type(c_ptr), value :: h
          integer(c_int32_t), value :: ns
          real(c_double), intent(inout) :: dt
          real(c_double), value :: dt_o
          real(c_double), intent(out) :: cflm
        end function fribo_cfl_dt

        integer(c_int) function fribo_substep(h, n, ns) bind(c, name="fribo_substep")
          import :: c_int, c_int32_t, c_ptr
          type(c_ptr), value :: h
          integer(c_int32_t), value :: n, ns
        end function fribo_substep

        integer(c_long) function fribo_array_size(h, name) bind(c, name="fribo_array_size")
          import :: c_long, c_char, c_ptr
          type(c_ptr), value :: h
          character(kind=c_char), intent(in) :: name(*)
        end function fribo_array_size

        integer(c_int) function fribo_get_array(h, name, out, n) bind(c, name="fribo_get_array")
          import :: c_int, c_long, c_char, c_ptr
          type(c_ptr), value :: h
          character(kind=c_char), intent(in) :: name(*)
          type(c_ptr), value :: out
          integer(c_long), value :: n
        end function fribo_get_array

        integer(c_int) function fribo_sync(h) bind(c, name="fribo_sync")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function fribo_sync

        ! loads_transfer_c (loads_transfer.F90:4-162) on the device
        integer(c_int) function fribo_loads_transfer(h, kprobe, lmnor, lmacc, press, tau) &
            bind(c, name="fribo_loads_transfer")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), value :: kprobe
          real(c_double), intent(in) :: lmnor(*), lmacc(*)
          real(c_double), intent(out) :: press(*), tau(*)
        end function fribo_loads_transfer

        ! checkbcs (ckbcs.F90:5-86) on the device
        integer(c_int) function fribo_checkbcs(h, bcerrv, bcerrw, nbcv, nbcw) bind(c, name="fribo_checkbcs")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(out) :: bcerrv(*), bcerrw(*)
          real(c_double), intent(out) :: nbcv, nbcw
        end function fribo_checkbcs

        ! axis of the Poisson eigen-transform: 0 z (reference order), 1 y, -1 auto
        integer(c_int) function fribo_set_transform(h, axis) bind(c, name="fribo_set_transform")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: h
          integer(c_int32_t), value :: axis
        end function fribo_set_transform

        ! mirror form of the transform along y on grids symmetric about the
        ! channel centre: 1 (the default where available), 0 the dense form
        integer(c_int) function fribo_set_ymirror(h, on) bind(c, name="fribo_set_ymirror")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: h
          integer(c_int32_t), value :: on
        end function fribo_set_ymirror

        integer(c_int) function fribo_diagnostics(h, out) bind(c, name="fribo_diagnostics")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: h
          real(c_double), intent(out) :: out(9)
        end function fribo_diagnostics

        integer(c_int) function fribo_fetch_fields(h, mask, kfirst, klast, q2, q3, pr, ru2, ru3) &
            bind(c, name="fribo_fetch_fields")
          import :: c_int, c_int32_t, c_double, c_ptr
          type(c_ptr), value :: h
          integer(c_int32_t), value :: mask, kfirst, klast
          real(c_double), intent(inout) :: q2(*), q3(*), pr(*), ru2(*), ru3(*)
        end function fribo_fetch_fields

        type(c_ptr) function fribo_last_error() bind(c, name="fribo_last_error")
          import :: c_ptr
        end function fribo_last_error

      end interface

      contains

       !message of the last failing call, as a Fortran string
        function fribo_error_message() result(msg)
          character(len=256) :: msg
          character(kind=c_char), pointer :: cs(:)
          type(c_ptr) :: p
          integer :: i
          msg = ' '
          p = fribo_last_error()
          if (.not. c_associated(p)) return
          call c_f_pointer(p, cs, [256])
          do i = 1, 256
            if (cs(i) == c_null_char) exit
            msg(i:i) = cs(i)
          enddo
        end function fribo_error_message

      END MODULE fribo_gpu_mod
