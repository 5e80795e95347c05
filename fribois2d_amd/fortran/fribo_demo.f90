!  Fortran driver of the GPU fluid step through fribo_gpu_mod.
!
!  Follows the time loop of the reference driver (main.F90:79-101) with
!  the fluid step replaced by fribo_fluid_tsch: a cylinder (D = 1.5) in a
!  uniform 48 x 36 channel, CFL-controlled dt, 20 steps.  On a uniform grid
!  the Neumann z operator of phini.F90:133-142 is dx3q times the 1-D
!  Neumann Laplacian, whose eigen-pairs are the DCT-II vectors
!  cos(pi*m*(k-1/2)/n3m) with eigenvalues -4*dx3q*sin(pi*m/(2*n3m))**2, so no
!  LAPACK is needed here.  After every step the host reads ttime's
!  diagnostics through the fribo_ttime shims (vmaxv, vminv, divgck, kinen,
!  getintvort, cfl_diag) instead of copying the fields back; at the end the
!  fields come back once, whole (fribo_get_state) and as a window of planes
!  (fribo_fetch_fields).  All inputs, the diagnostics and the final state are
!  written to fribo_demo_out.bin (stream access) for comparison with other
!  drivers.

      PROGRAM fribo_demo

      USE iso_c_binding
      USE fribo_gpu_mod
      USE fribo_ttime

      implicit none

      integer, parameter :: n2 = 37, n3 = 49, n2m = n2-1, n3m = n3-1
      integer, parameter :: nlm = 40, nsteps = 20
      real(c_double), parameter :: pi = 3.14159265358979323846d0
      type(fribo_params) :: p
      type(c_ptr) :: h
      real(c_double) :: rc(n2), zz(n3), wr(n3m), zmm(n3m,n3m), zmmt(n3m,n3m)
      real(c_double) :: q2(n2,-2:n3m+3), q3(n2,-2:n3m+3), pr(n2,-2:n3m+3)
      real(c_double) :: ru2(n2,n3m), ru3(n2,n3m)
      real(c_double) :: diag(9,nsteps), w2(n2,-2:n3m+3), w3(n2,-2:n3m+3), wp(n2,-2:n3m+3)
      real(c_double) :: wru2(n2,n3m), wru3(n2,n3m), vmx(2), vmn(2), dmx, ek, vo, cfd, q3a
      real(c_double) :: qb2s(n2), qb3s(n2), qb2n(n2), qb3n(n2), dq2x2o(n2), dq3x2o(n2)
      real(c_double) :: vert(nlm+1,2), lmcc(nlm,2), lmvel(nlm,2), lmA(nlm)
      real(c_double) :: dt, dt_o, cflm, dx3q, th, nrm, dthist(nsteps)
      integer(c_int) :: rc_, ns
      integer :: j, k, m

      p%n2 = n2
      p%n3 = n3
      p%nsst = 3
      p%idtv = 0
      p%ibact = 1
      p%forst = 1
      p%wfun = 1
      p%inslws = 0
      p%inslwn = 0
      p%alx3 = 8.0d0
      p%rext2 = 6.0d0
      p%visc = 0.02d0
      p%cou = 0.8d0
      p%q3uw = 0.0d0
      p%q3lw = 0.0d0
      p%dtmax = 0.01d0
      p%cflmax = 0.3d0
      p%cfllim = 2.0d0
      p%radiusSD = 1.6d0

      do j = 1, n2
        rc(j) = p%rext2*real(j-1,c_double)/real(n2m,c_double)
      enddo
      do k = 1, n3
        zz(k) = p%alx3*real(k-1,c_double)/real(n3m,c_double)
      enddo

     !eigen-pairs of the Neumann z operator (uniform grid)
      dx3q = (real(n3m,c_double)/p%alx3)**2
      do m = 1, n3m
        th = pi*real(m-1,c_double)/real(n3m,c_double)
        wr(m) = -4.0d0*dx3q*sin(0.5d0*th)**2
        nrm = sqrt(real(n3m,c_double))
        if (m > 1) nrm = sqrt(0.5d0*real(n3m,c_double))
        do k = 1, n3m
          zmm(k,m) = cos(th*(real(k,c_double)-0.5d0))/nrm
        enddo
      enddo
      zmmt = transpose(zmm)

     !initial solution as inqpr (initial_sol.F90:121-149)
      q2 = 0.0d0
      q3 = 0.0d0
      pr = 0.0d0
      q3(:,0:n3) = 1.0d0
      ru2 = 0.0d0
      ru3 = 0.0d0
      qb2s = 0.0d0
      qb2n = 0.0d0
      qb3s = 0.0d0
      qb3n = 0.0d0
      qb3s(1:n2m) = 1.0d0
      qb3n(1:n2m) = 1.0d0
      dq2x2o = 0.0d0
      dq3x2o = 0.0d0

     !closed polygon and markers as read_rigid (utils_proc.F90:560-565)
      do m = 1, nlm+1
        th = 2.0d0*pi*real(m-1,c_double)/real(nlm,c_double)
        vert(m,1) = 3.0d0 + 0.75d0*sin(th)
        vert(m,2) = 3.0d0 + 0.75d0*cos(th)
      enddo
      do m = 1, nlm
        lmcc(m,1) = (vert(m,1)+vert(m+1,1))*0.5d0
        lmcc(m,2) = (vert(m,2)+vert(m+1,2))*0.5d0
        lmA(m) = sqrt((vert(m+1,1)-vert(m,1))**2+(vert(m+1,2)-vert(m,2))**2)
      enddo
      lmvel = 0.0d0

      rc_ = fribo_init(p, rc, zz, zmm, zmmt, int(n3m,c_int), wr, h)
      call check(rc_, 'fribo_init')
      fribo_diag_h = h
      rc_ = fribo_set_state(h, q2, q3, pr, ru2, ru3, qb2s, qb3s, qb2n, qb3n, dq2x2o, dq3x2o)
      call check(rc_, 'fribo_set_state')

      dt = 1.0d-3
      dt_o = dt
      do ns = 0, nsteps-1
        rc_ = fribo_set_markers(h, int(nlm,c_int32_t), lmcc, lmvel, lmA)
        call check(rc_, 'fribo_set_markers')
        rc_ = fribo_fluid_tsch(h, int(ns,c_int32_t), dt, dt_o, cflm)
        call check(rc_, 'fribo_fluid_tsch')
        dthist(ns+1) = dt
        dt_o = dt
       !ttime's diagnostics of this step from the device
        fribo_diag_valid = .false.
        call vmaxv(vmx)
        call vminv(vmn)
        call cfl_diag(cfd, q3a)
        call divgck(dmx)
        call getintvort(vo)
        call kinen(ek)
        diag(:,ns+1) = (/ vmx(1), vmx(2), vmn(1), vmn(2), dmx, ek, vo, cfd, q3a /)
      enddo

      rc_ = fribo_get_state(h, q2, q3, pr, ru2, ru3, qb2s, qb3s, qb2n, qb3n, dq2x2o, dq3x2o)
      call check(rc_, 'fribo_get_state')
     !planes 5..17 of every field through the windowed read-back
      w2 = 0.0d0
      w3 = 0.0d0
      wp = 0.0d0
      wru2 = 0.0d0
      wru3 = 0.0d0
      rc_ = fribo_fetch_fields(h, 31_c_int32_t, 5_c_int32_t, 17_c_int32_t, w2, w3, wp, wru2, wru3)
      call check(rc_, 'fribo_fetch_fields')
      if (any(w2(:,5:17) /= q2(:,5:17)) .or. any(w3(:,5:17) /= q3(:,5:17)) .or. &
          any(wp(:,5:17) /= pr(:,5:17)) .or. any(wru2(:,5:17) /= ru2(:,5:17)) .or. &
          any(wru3(:,5:17) /= ru3(:,5:17)) .or. any(w2(:,-2:4) /= 0.0d0) .or. any(wp(:,18:) /= 0.0d0)) then
        write(*,*) 'fribo_fetch_fields: window differs from fribo_get_state'
        stop 1
      endif
      rc_ = fribo_finalize(h)

      open(21, file='fribo_demo_out.bin', access='stream', form='unformatted', status='replace')
      write(21) n2, n3, nlm, nsteps
      write(21) rc, zz, zmm, zmmt, wr, lmcc, lmvel, lmA, dthist, q2, q3, pr, diag
      close(21)

      write(*,'(a,i3,a,es12.5,a,es12.5)') ' fribo_demo: ', nsteps, ' steps, last dt =', dt, &
        ', max|q2| =', maxval(abs(q2))

      contains

        subroutine check(code, what)
          integer(c_int), intent(in) :: code
          character(*), intent(in) :: what
          if (code /= FRIBO_OK) then
            write(*,*) what, ' failed with code ', code, ': ', trim(fribo_error_message())
            stop 1
          endif
        end subroutine check

      END PROGRAM fribo_demo
