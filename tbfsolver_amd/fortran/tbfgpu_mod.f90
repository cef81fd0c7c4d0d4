! tbfgpu_mod.f90 -- Fortran 2003 iso_c_binding interface to libtbfgpu.so
! (the C ABI in include/tbfgpu.h), plus wrappers with the names and argument
! meaning of the reference's pressure-solver plug-ins:
!   tbfg_fastPoissonSolverCTOR / tbfg_solveFPS  <->  fastPoissonSolverCTOR / solveFPS
!        (src/fastPoissonSolver/fastPoissonSolver.f90:123-156, 337-384)
!   tbfg_pcgCTOR / tbfg_solvePCG                <->  pcgCTOR / solvePCG
!        (src/pcg/pcg.f90:62-99, 103-140)
! Field arrays are passed with their own bounds (any halo); the interior 1..n is
! read and the output halo is filled as updateBoundaries does for the channel.
module tbfgpu_mod
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: TBFG_OK = 0, TBFG_EINVAL = 1, TBFG_ENOMEM = 2, &
                                 TBFG_ENOTCONVERGED = 3, TBFG_ENUMERIC = 4, TBFG_EDEVICE = 5
    integer(c_int32_t), parameter :: TBFG_PREC_FP64 = 0, TBFG_PREC_MIXED = 1, TBFG_PREC_MIXED_BF16 = 2

    type, bind(C) :: tbfg_grid_desc
        integer(c_int32_t) :: n(3)
        real(c_double) :: L(3)
        integer(c_int32_t) :: stretch(3)
        real(c_double) :: gamma
    end type

    type, bind(C) :: tbfg_fps_opts
        integer(c_int32_t) :: precision
        integer(c_int32_t) :: refine_sweeps
    end type

    type, bind(C) :: tbfg_pcg_params
        integer(c_int32_t) :: n_levels, n_pre, n_post, max_iter
        real(c_double) :: tol
        integer(c_int32_t) :: mg_cycles, precision, coarse_max_iter, flexible
        real(c_double) :: coarse_tol
        integer(c_int32_t) :: preconditioner, mg_max_iter
        real(c_double) :: mg_tol
    end type

    interface
        function tbfg_last_error_c() bind(C, name='tbfg_last_error') result(p)
            import
            type(c_ptr) :: p
        end function

        integer(c_int) function tbfg_fps_create(grid, opts, h) bind(C, name='tbfg_fps_create')
            import
            type(tbfg_grid_desc), intent(in) :: grid
            type(tbfg_fps_opts), intent(in) :: opts
            type(c_ptr), intent(out) :: h
        end function

        integer(c_int) function tbfg_fps_solve(h, s, s_lb, s_ub, p, p_lb, p_ub) bind(C, name='tbfg_fps_solve')
            import
            type(c_ptr), value :: h
            real(c_double), intent(in) :: s(*)
            integer(c_int64_t), intent(in) :: s_lb(3), s_ub(3)
            real(c_double), intent(inout) :: p(*)
            integer(c_int64_t), intent(in) :: p_lb(3), p_ub(3)
        end function

        integer(c_int) function tbfg_fps_destroy(h) bind(C, name='tbfg_fps_destroy')
            import
            type(c_ptr), value :: h
        end function

        subroutine tbfg_pcg_default_params(p) bind(C, name='tbfg_pcg_default_params')
            import
            type(tbfg_pcg_params), intent(out) :: p
        end subroutine

        integer(c_int) function tbfg_pcg_create(grid, prm, h) bind(C, name='tbfg_pcg_create')
            import
            type(tbfg_grid_desc), intent(in) :: grid
            type(tbfg_pcg_params), intent(in) :: prm
            type(c_ptr), intent(out) :: h
        end function

        integer(c_int) function tbfg_pcg_solve(h, psi, psi_lb, psi_ub, beta, beta_lb, beta_ub, q, q_lb, q_ub, &
                                               iters, rel_res) bind(C, name='tbfg_pcg_solve')
            import
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: psi(*)
            integer(c_int64_t), intent(in) :: psi_lb(3), psi_ub(3)
            real(c_double), intent(in) :: beta(*)
            integer(c_int64_t), intent(in) :: beta_lb(3), beta_ub(3)
            real(c_double), intent(in) :: q(*)
            integer(c_int64_t), intent(in) :: q_lb(3), q_ub(3)
            integer(c_int32_t), intent(out) :: iters
            real(c_double), intent(out) :: rel_res
        end function

        integer(c_int) function tbfg_pcg_destroy(h) bind(C, name='tbfg_pcg_destroy')
            import
            type(c_ptr), value :: h
        end function

        ! FAST_MODE step on device arrays (computeSource + computeOldPressDiv,
        ! poissonEqn.f90:193-305; makeVelocityDivFree + computeOldPressGrad,
        ! momentumEqn.f90:999-1127; updatePressure + clean_st_prev_bc,
        ! poissonEqn.f90:374-459).  Every field argument is a device address
        ! (c_ptr).  st, st1, st2 are c_loc of a type(c_ptr), target :: st(3) array
        ! holding the x, y, z components' device addresses, or c_null_ptr.
        integer(c_int) function tbfg_source_fast_dev(grid, alpha_dt, rho0, rho, ux, uy, uz, psi1, psi2, &
                                                     st1, st2, s, stream) bind(C, name='tbfg_source_fast_dev')
            import
            type(tbfg_grid_desc), intent(in) :: grid
            real(c_double), value :: alpha_dt, rho0
            type(c_ptr), value :: rho, ux, uy, uz, psi1, psi2, st1, st2, s, stream
        end function

        integer(c_int) function tbfg_project_fast_dev(grid, alpha_dt, rho0, rho, psi, psi1, psi2, st1, st2, &
                                                      ux, uy, uz, stream) bind(C, name='tbfg_project_fast_dev')
            import
            type(tbfg_grid_desc), intent(in) :: grid
            real(c_double), value :: alpha_dt, rho0
            type(c_ptr), value :: rho, psi, psi1, psi2, st1, st2, ux, uy, uz, stream
        end function

        integer(c_int) function tbfg_update_pressure_dev(grid, psi, p, psi1, psi2, st, st1, st2, stream) &
                bind(C, name='tbfg_update_pressure_dev')
            import
            type(tbfg_grid_desc), intent(in) :: grid
            type(c_ptr), value :: psi, p, psi1, psi2, st, st1, st2, stream
        end function
    end interface

contains

    ! message of the calling thread's last failure
    function tbfg_last_error() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: cs(:)
        type(c_ptr) :: p
        integer :: n
        p = tbfg_last_error_c()
        call c_f_pointer(p, cs, [4096])
        n = 0
        do while (n < 4096)
            if (cs(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(cs(1:n), msg)
    end function

    ! channel grid: x, z periodic uniform; y walls, uniform (stretch_y = .false.)
    ! or tanh-stretched with factor gamma (grid.f90:293-330)
    function tbfg_channel_grid(nx, ny, nz, Lx, Ly, Lz, stretch_y, gamma) result(g)
        integer, intent(in) :: nx, ny, nz
        real(c_double), intent(in) :: Lx, Ly, Lz, gamma
        logical, intent(in) :: stretch_y
        type(tbfg_grid_desc) :: g
        g%n = [int(nx, c_int32_t), int(ny, c_int32_t), int(nz, c_int32_t)]
        g%L = [Lx, Ly, Lz]
        g%stretch = 0
        if (stretch_y) g%stretch(2) = 1
        g%gamma = gamma
    end function

    ! fastPoissonSolverCTOR
    subroutine tbfg_fastPoissonSolverCTOR(h, grid, mixed, ierr)
        type(c_ptr), intent(out) :: h
        type(tbfg_grid_desc), intent(in) :: grid
        logical, intent(in) :: mixed
        integer, intent(out) :: ierr
        type(tbfg_fps_opts) :: o
        o%precision = TBFG_PREC_FP64
        if (mixed) o%precision = TBFG_PREC_MIXED
        o%refine_sweeps = 2
        ierr = tbfg_fps_create(grid, o, h)
    end subroutine

    ! solveFPS(this, p, s): p with halo (inout), s (in)
    subroutine tbfg_solveFPS(h, p, s, ierr)
        type(c_ptr), intent(in) :: h
        real(c_double), allocatable, intent(inout) :: p(:,:,:)
        real(c_double), allocatable, intent(in) :: s(:,:,:)
        integer, intent(out) :: ierr
        integer(c_int64_t) :: plb(3), pub(3), slb(3), sub(3)
        plb = lbound(p); pub = ubound(p)
        slb = lbound(s); sub = ubound(s)
        ierr = tbfg_fps_solve(h, s, slb, sub, p, plb, pub)
    end subroutine

    ! pcgCTOR: parameters with the keys of specs/pcg_solver.  The optional
    ! flexible selects the direction update: .false. Fletcher-Reeves as
    ! pcg.f90:241-255 (the FP64 default), .true. Polak-Ribiere, which keeps
    ! converging where the V-cycle preconditioner (red-black order in both pre-
    ! and post-smoothing) is not symmetric enough for Fletcher-Reeves.  The
    ! optional tolMG and maxIterMG select solveMG's loop (V-cycles per
    ! application until the finest residual is at most tolMG); without them
    ! one V-cycle per application.
    subroutine tbfg_pcgCTOR(h, grid, nLevels, nPreSweep, nPostSweep, tolPCG, maxIterPCG, mixed, ierr, flexible, &
                            tolMG, maxIterMG)
        type(c_ptr), intent(out) :: h
        type(tbfg_grid_desc), intent(in) :: grid
        integer, intent(in) :: nLevels, nPreSweep, nPostSweep, maxIterPCG
        real(c_double), intent(in) :: tolPCG
        logical, intent(in) :: mixed
        integer, intent(out) :: ierr
        logical, intent(in), optional :: flexible
        real(c_double), intent(in), optional :: tolMG
        integer, intent(in), optional :: maxIterMG
        type(tbfg_pcg_params) :: prm
        call tbfg_pcg_default_params(prm)
        prm%n_levels = nLevels
        prm%n_pre = nPreSweep
        prm%n_post = nPostSweep
        prm%tol = tolPCG
        prm%max_iter = maxIterPCG
        if (mixed) prm%precision = TBFG_PREC_MIXED
        if (present(flexible)) then
            prm%flexible = 0
            if (flexible) prm%flexible = 1
        end if
        if (present(tolMG)) prm%mg_tol = tolMG
        if (present(maxIterMG)) prm%mg_max_iter = maxIterMG
        ierr = tbfg_pcg_create(grid, prm, h)
    end subroutine

    ! solvePCG(this, mesh, psi, beta, q); iter_ and res_ as in pcg.f90:36-37
    subroutine tbfg_solvePCG(h, psi, beta, q, iter_, res_, ierr)
        type(c_ptr), intent(in) :: h
        real(c_double), allocatable, intent(inout) :: psi(:,:,:)
        real(c_double), allocatable, intent(in) :: beta(:,:,:), q(:,:,:)
        integer, intent(out) :: iter_
        real(c_double), intent(out) :: res_
        integer, intent(out) :: ierr
        integer(c_int64_t) :: a(3), b(3), c(3), d(3), e(3), f(3)
        integer(c_int32_t) :: it
        a = lbound(psi); b = ubound(psi)
        c = lbound(beta); d = ubound(beta)
        e = lbound(q); f = ubound(q)
        ierr = tbfg_pcg_solve(h, psi, a, b, beta, c, d, q, e, f, it, res_)
        iter_ = it
        if (ierr == TBFG_OK) iter_ = it + 1   ! the reference counts the final stop test
    end subroutine

end module tbfgpu_mod
