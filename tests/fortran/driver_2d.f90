! Driver written against robinsolver's public Fortran API (the usage of
! example.f90:1-89 and README.md:36-126): 2D unit square, k = kb = 1,
! manufactured solution cos(x/sqrt 2) cos(y/sqrt 2) with its impedance data,
! f = 0.  Prints the relative l2 error of the leaf solutions; with a fourth
! argument, writes every leaf's solution vector (DFS leaf order, complex128,
! stream access) to that file.
program driver_2d
    use derived_type_module
    use local_solver_module
    use robin_tree_module
    use robin_output_module
    implicit none
    integer, parameter :: dp = kind(1.0d0)
    type(elliptic_operator) :: op
    type(solve_opts) :: opts
    type(robin_tree) :: tree
    type(elliptic_rhs) :: rhs
    type(leaf_solution), dimension(:), allocatable :: u_out, zero_sol
    real(dp), dimension(:), allocatable :: errs, nrm
    character(len=32) :: arg
    character(len=512) :: outpath
    integer :: n, L, q, unit_out
    real(dp) :: num, den

    n = 64
    L = 4
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) n
    end if
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) L
    end if
    if (command_argument_count() >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) opts%precision
    end if
    op%d = 2
    op%domain(:,1) = [0.0d0, 1.0d0]
    op%domain(:,2) = [0.0d0, 1.0d0]
    op%domain(:,3) = [0.0d0, 1.0d0]
    op%k => wavespeed
    opts%h_tgt = 1.0d0 / real(n, dp)
    opts%kb = 1.0d0
    tree%bndry_params(1,:) = [cmplx(0.0,1.0,dp)*opts%kb, cmplx(1.0,0.0,dp)]
    tree%bndry_params(2,:) = [cmplx(0.0,-1.0,dp)*opts%kb, cmplx(1.0,0.0,dp)]
    rhs%d = 2
    rhs%f => zero_f
    rhs%g(1)%g_face => lo_face
    rhs%g(2)%g_face => hi_face
    rhs%g(3)%g_face => lo_face
    rhs%g(4)%g_face => hi_face

    call tree%factor(L, op, opts)
    call tree%solve(rhs)
    call tree%solve(rhs)
    call get_solution(tree, u_out)
    call check_solution(tree%ell_op%d, u_out, exact, tree%opts, errs, '2  ', .false.)
    allocate(zero_sol(size(u_out)))
    do q = 1, size(u_out)
        zero_sol(q)%box = u_out(q)%box
        allocate(zero_sol(q)%u(size(u_out(q)%u)))
        zero_sol(q)%u = (0.0d0, 0.0d0)
    end do
    call check_solution(tree%ell_op%d, zero_sol, exact, tree%opts, nrm, '2  ', .false.)
    num = sqrt(sum(errs**2))
    den = sqrt(sum(nrm**2))
    print '(A,ES12.4)', 'RELL2 ', num / den
    if (command_argument_count() >= 4) then
        call get_command_argument(4, outpath)
        open(newunit=unit_out, file=trim(outpath), access='stream', form='unformatted', status='replace')
        do q = 1, size(u_out)
            write(unit_out) u_out(q)%u
        end do
        close(unit_out)
    end if
    call tree%inspect()

contains

    pure function wavespeed(x) result(r)
        real(dp), dimension(:), intent(in) :: x
        complex(dp) :: r
        r = (1.0d0, 0.0d0)
    end function wavespeed

    pure function zero_f(x) result(r)
        real(dp), dimension(:), intent(in) :: x
        complex(dp) :: r
        r = (0.0d0, 0.0d0)
    end function zero_f

    ! impedance data i u - du/dx at x = 0 (and the y analogue); t is the tangential coordinate
    pure function lo_face(x) result(r)
        real(dp), dimension(:), intent(in) :: x
        complex(dp) :: r
        r = cmplx(0.0d0, 1.0d0, dp) * cos(x(1) / sqrt(2.0d0))
    end function lo_face

    pure function hi_face(x) result(r)
        real(dp), dimension(:), intent(in) :: x
        complex(dp) :: r
        r = cmplx(0.0d0, 1.0d0, dp) * cos(1.0d0 / sqrt(2.0d0)) * cos(x(1) / sqrt(2.0d0)) &
            - 1.0d0 / sqrt(2.0d0) * sin(1.0d0 / sqrt(2.0d0)) * cos(x(1) / sqrt(2.0d0))
    end function hi_face

    pure function exact(x) result(r)
        real(dp), dimension(1:3), intent(in) :: x
        complex(dp) :: r
        r = cos(x(1) / sqrt(2.0d0)) * cos(x(2) / sqrt(2.0d0))
    end function exact


end program driver_2d
