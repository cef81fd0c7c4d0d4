! tbfgpu_test.f90 -- Fortran driver exercising libtbfgpu.so through tbfgpu_mod
! the way the reference's poissonEqn would: fields with one halo layer
! (lower bound 0), solveFPS and solvePCG on a 32^3 channel.
!
! Right-hand side: s = cos(2 pi x / Lx) (uniform in y and z).  The discrete
! solution of both operators (FFT operator; 7-point operator with beta = 1) is
! s / lambda with lambda = 2 (cos(2 pi / nx) - 1) / dx^2, which is mean free.
! On this single-mode right-hand side Fletcher-Reeves PCG stagnates (as the
! reference's does: its V-cycle smooths red-black before and after the coarse
! correction), so the PCG case uses the Polak-Ribiere update, and the
! Fletcher-Reeves call checks that stagnation is reported as not converged.
program tbfgpu_test
    use tbfgpu_mod
    implicit none
    integer, parameter :: n = 32
    real(c_double), parameter :: pi = 3.141592653589793d0
    real(c_double), allocatable :: s(:,:,:), p(:,:,:), beta(:,:,:), psi(:,:,:)
    type(tbfg_grid_desc) :: g
    type(c_ptr) :: fps, pcg
    integer :: ierr, i, iter_
    real(c_double) :: Lx, dx, lam, x, err, res_

    Lx = 2.d0 * pi
    dx = Lx / n
    lam = 2.d0 * (cos(2.d0 * pi / n) - 1.d0) / (dx * dx)
    g = tbfg_channel_grid(n, n, n, Lx, 2.d0, pi, .true., 1.5d0)

    allocate(s(0:n+1, 0:n+1, 0:n+1), p(0:n+1, 0:n+1, 0:n+1))
    allocate(beta(0:n+1, 0:n+1, 0:n+1), psi(0:n+1, 0:n+1, 0:n+1))
    do i = 1, n
        x = (i - 0.5d0) * dx
        s(i, :, :) = cos(2.d0 * pi * x / Lx)
    end do
    p = 0.d0

    call tbfg_fastPoissonSolverCTOR(fps, g, .false., ierr)
    if (ierr /= TBFG_OK) call fail('fastPoissonSolverCTOR: ' // tbfg_last_error())
    call tbfg_solveFPS(fps, p, s, ierr)
    if (ierr /= TBFG_OK) call fail('solveFPS: ' // tbfg_last_error())
    err = maxval(abs(p(1:n, 1:n, 1:n) - s(1:n, 1:n, 1:n) / lam))
    if (err > 1d-11) call fail('solveFPS error too large')
    if (any(p(0, 1:n, 1:n) /= p(n, 1:n, 1:n))) call fail('solveFPS: periodic halo')
    if (any(p(1:n, 0, 1:n) /= p(1:n, 1, 1:n))) call fail('solveFPS: Neumann halo')
    print '(a, es10.3)', 'solveFPS max error ', err

    beta = 1.d0
    psi = 0.d0
    call tbfg_pcgCTOR(pcg, g, 4, 2, 2, 1.d-12, 100, .false., ierr, flexible = .true.)
    if (ierr /= TBFG_OK) call fail('pcgCTOR: ' // tbfg_last_error())
    call tbfg_solvePCG(pcg, psi, beta, s, iter_, res_, ierr)
    if (ierr /= TBFG_OK) call fail('solvePCG: ' // tbfg_last_error())
    psi = psi - sum(psi(1:n, 1:n, 1:n)) / real(n**3, c_double)
    err = maxval(abs(psi(1:n, 1:n, 1:n) - s(1:n, 1:n, 1:n) / lam))
    if (err > 1d-9) call fail('solvePCG error too large')
    print '(a, i4, a, es10.3, a, es10.3)', 'solvePCG iter_ ', iter_, ' res_ ', res_, ' max error ', err

    ierr = tbfg_fps_destroy(fps)
    ierr = tbfg_pcg_destroy(pcg)

    psi = 0.d0
    call tbfg_pcgCTOR(pcg, g, 4, 2, 2, 1.d-12, 100, .false., ierr)
    if (ierr /= TBFG_OK) call fail('pcgCTOR (Fletcher-Reeves): ' // tbfg_last_error())
    call tbfg_solvePCG(pcg, psi, beta, s, iter_, res_, ierr)
    if (ierr /= TBFG_ENOTCONVERGED .or. iter_ /= 100) call fail('expected TBFG_ENOTCONVERGED after 100')
    print '(a, i4, a, es10.3)', 'solvePCG (Fletcher-Reeves) not converged: iter_ ', iter_, ' res_ ', res_
    ierr = tbfg_pcg_destroy(pcg)

    ! solveMG's loop with the reference channel case's tolMG and maxIterMG
    psi = 0.d0
    call tbfg_pcgCTOR(pcg, g, 4, 2, 2, 1.d-12, 100, .false., ierr, flexible = .true., tolMG = 0.4d0, &
                      maxIterMG = 100)
    if (ierr /= TBFG_OK) call fail('pcgCTOR (tolMG): ' // tbfg_last_error())
    call tbfg_solvePCG(pcg, psi, beta, s, iter_, res_, ierr)
    if (ierr /= TBFG_OK) call fail('solvePCG (tolMG): ' // tbfg_last_error())
    psi = psi - sum(psi(1:n, 1:n, 1:n)) / real(n**3, c_double)
    err = maxval(abs(psi(1:n, 1:n, 1:n) - s(1:n, 1:n, 1:n) / lam))
    if (err > 1d-9) call fail('solvePCG (tolMG) error too large')
    print '(a, i4, a, es10.3)', 'solvePCG (tolMG 0.4) iter_ ', iter_, ' max error ', err
    ierr = tbfg_pcg_destroy(pcg)

    ! error path: non-divisible hierarchy is rejected with a status code
    call tbfg_pcgCTOR(pcg, tbfg_channel_grid(36, 32, 32, Lx, 2.d0, pi, .false., 1.4d0), 4, 2, 2, &
                      1.d-8, 100, .false., ierr)
    if (ierr /= TBFG_EINVAL) call fail('expected TBFG_EINVAL')
    print '(a)', 'tbfgpu_test OK'

contains
    subroutine fail(msg)
        character(len=*), intent(in) :: msg
        print '(a)', 'FAIL: ' // msg
        stop 1
    end subroutine
end program tbfgpu_test
