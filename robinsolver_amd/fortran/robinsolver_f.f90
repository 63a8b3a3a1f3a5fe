! =================================================================================================
! robinsolver_f.f90 -- Fortran drop-in for robinsolver's public API, backed by the MI355X library.
!
! Keeps the reference's module, type and procedure names so that existing drivers recompile and
! link unchanged:
!   derived_type_module   elliptic_operator, elliptic_rhs, bndry_rhs, solve_opts, debug_opts,
!                         bndry_operator, bndry_vector, coeff_proc, scalar_proc, sol_proc
!                         (derived_type_module.f90:16-320)
!   time_module           start_clock, read_clock (time_module.f90:14-36)
!   local_solver_module   present so that `use local_solver_module` compiles; drivers call
!                         nothing from it (local_solver_module.f90:28-29)
!   robin_tree_module     type robin_tree with factor, solve, memsize, inspect
!                         (robin_tree_module.f90:55-274)
!   robin_output_module   leaf_solution, get_solution, check_solution, write_variable
!                         (robin_output_module.f90:19-138)
! The tree itself lives on the GPU behind an opaque handle (include/robinsolver.h).  The callbacks
! k, f and g_face are evaluated here, on the host, at the points and in the order the reference
! evaluates them.  Two components are added to solve_opts with defaults that keep the reference
! behaviour: precision (0 = FP64 path, 1 = mixed path) and n_refine.
! =================================================================================================

module derived_type_module
    use iso_c_binding
    implicit none
    private
    integer, parameter :: dp = kind(1.0d0)

    public :: coeff_proc, scalar_proc, sol_proc, elliptic_operator, elliptic_rhs, bndry_rhs, &
              bndry_operator, bndry_vector, solve_opts, debug_opts

    abstract interface
        pure function coeff_proc(x) result(c)
            import dp
            real(dp), dimension(:), intent(in) :: x
            complex(dp), dimension(1:3,1:3) :: c
        end function
        pure function scalar_proc(x) result(r)
            import dp
            real(dp), dimension(:), intent(in) :: x
            complex(dp) :: r
        end function
        pure function sol_proc(x) result(r)
            import dp
            real(dp), dimension(1:3), intent(in) :: x
            complex(dp) :: r
        end function
    end interface

    type elliptic_operator
        integer :: d
        real(dp), dimension(1:2,1:3) :: domain
        procedure(coeff_proc), pointer, nopass :: coeff => NULL()
        procedure(scalar_proc), pointer, nopass :: k => NULL()
        complex(dp), dimension(1:6,1:2) :: robin_coeff
    end type elliptic_operator

    type bndry_rhs
        procedure(scalar_proc), pointer, nopass :: g_face => NULL()
    end type bndry_rhs

    type elliptic_rhs
        integer :: d
        procedure(scalar_proc), pointer, nopass :: f
        type(bndry_rhs), dimension(1:6) :: g
    end type elliptic_rhs

    type bndry_operator
        complex(dp), allocatable, dimension(:,:) :: mat
    end type bndry_operator

    type bndry_vector
        complex(dp), allocatable, dimension(:) :: vec
    end type bndry_vector

    type debug_opts
        logical :: txt_output = .false.
        logical :: no_delete = .false.
    end type debug_opts

    type solve_opts
        real(dp) :: h_tgt
        real(dp), dimension(1:3) :: h
        integer :: ordering = 1
        character(len=2) :: disc = 'fd'
        character(len=3) :: local_mode = 'LU'
        character(len=3) :: R_mode = 'LU'
        real(dp) :: kb = 1.0d0
        type(debug_opts) :: debug
        integer :: precision = 0
        integer :: n_refine = 2
    end type solve_opts

end module derived_type_module


module time_module
    implicit none
    private
    integer, parameter :: dp = kind(1.0d0)
    integer(8) :: clock0, clock1, clockrate
    public :: start_clock, read_clock
contains
    subroutine start_clock()
        call system_clock(clock0, count_rate=clockrate)
    end subroutine start_clock

    subroutine read_clock(time)
        real(dp), intent(out) :: time
        call system_clock(clock1, count_rate=clockrate)
        time = real(clock1 - clock0, dp) / real(clockrate, dp)
        write(*, '(A,ES10.2)') 'Seconds elapsed: ', time
    end subroutine read_clock
end module time_module


module local_solver_module
    implicit none
    private
end module local_solver_module


module rs_capi
    ! ISO_C_BINDING interfaces of include/robinsolver.h
    use iso_c_binding
    implicit none
    public

    type, bind(C) :: rs_opts_t
        integer(c_int) :: precision
        integer(c_int) :: skip_root_rtr
        integer(c_int) :: n_refine
        integer(c_int) :: verbose
        integer(c_long_long) :: hbm_budget
        integer(c_int) :: autotune
        integer(c_int) :: graphs
        integer(c_int) :: merge_fp64_levels
        integer(c_int) :: retain
    end type rs_opts_t

    interface
        subroutine rs_opts_default(o) bind(C, name='rs_opts_default')
            import :: rs_opts_t
            type(rs_opts_t), intent(out) :: o
        end subroutine
        integer(c_int) function rs_tree_create(t, o) bind(C, name='rs_tree_create')
            import :: c_int, c_ptr, rs_opts_t
            type(c_ptr), intent(out) :: t
            type(rs_opts_t), intent(in) :: o
        end function
        integer(c_int) function rs_tree_destroy(t) bind(C, name='rs_tree_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: t
        end function
        integer(c_int) function rs_plan(t, d, L, domain, h_tgt, kb, bp, nl, m, ni, nb) bind(C, name='rs_plan')
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: t
            integer(c_int), value :: d, L
            real(c_double), intent(in) :: domain(*)
            real(c_double), value :: h_tgt, kb
            complex(c_double_complex), intent(in) :: bp(*)
            integer(c_int), intent(out) :: nl, m, ni, nb
        end function
        integer(c_int) function rs_leaf_points(t, leaf, xyz, face) bind(C, name='rs_leaf_points')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: t
            integer(c_int), value :: leaf
            real(c_double), intent(out) :: xyz(*)
            integer(c_int), intent(out) :: face(*)
        end function
        integer(c_int) function rs_leaf_box(t, leaf, box) bind(C, name='rs_leaf_box')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: t
            integer(c_int), value :: leaf
            real(c_double), intent(out) :: box(*)
        end function
        integer(c_int) function rs_leaf_grid(t, ptbox, h) bind(C, name='rs_leaf_grid')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: t
            integer(c_int), intent(out) :: ptbox(*)
            real(c_double), intent(out) :: h(*)
        end function
        integer(c_int) function rs_factor(t, k_int) bind(C, name='rs_factor')
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: t
            complex(c_double_complex), intent(in) :: k_int(*)
        end function
        integer(c_int) function rs_solve(t, nrhs, f, g, prec) bind(C, name='rs_solve')
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: t
            integer(c_int), value :: nrhs, prec
            complex(c_double_complex), intent(in) :: f(*), g(*)
        end function
        integer(c_int) function rs_get_solution(t, rhs, u) bind(C, name='rs_get_solution')
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: t
            integer(c_int), value :: rhs
            complex(c_double_complex), intent(out) :: u(*)
        end function
        type(c_ptr) function rs_last_error(t) bind(C, name='rs_last_error')
            import :: c_ptr
            type(c_ptr), value :: t
        end function
    end interface

contains

    function c_message(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=512) :: s
        character(kind=c_char), pointer :: cs(:)
        integer :: j
        s = ' '
        if (.not. c_associated(p)) return
        call c_f_pointer(p, cs, [512])
        do j = 1, 512
            if (cs(j) == c_null_char) exit
            s(j:j) = cs(j)
        end do
    end function c_message

    subroutine rs_check(rc, t)
        integer(c_int), intent(in) :: rc
        type(c_ptr), intent(in) :: t
        if (rc /= 0) then
            write(*, '(A)') trim(c_message(rs_last_error(t)))
            stop 1
        end if
    end subroutine rs_check

end module rs_capi


module robin_tree_module
    use iso_c_binding
    use derived_type_module
    use time_module
    use rs_capi
    implicit none
    private
    integer, parameter :: dp = kind(1.0d0)

    public :: robin_tree

    type robin_tree
        integer :: nlvls
        type(elliptic_operator) :: ell_op
        type(solve_opts) :: opts
        complex(dp), dimension(1:2,1:2) :: bndry_params
        logical :: factored = .false.
        real(dp) :: f_time
        real(dp) :: s_time
        type(c_ptr) :: handle = c_null_ptr
        integer :: n_leaves = 0, m = 0, n_int = 0, n_bnd = 0
    contains
        procedure :: factor => robin_tree_factor
        procedure :: solve => robin_tree_solve
        procedure :: memsize => robin_tree_memsize
        procedure :: inspect => robin_tree_inspect
        procedure :: leaf_box => robin_tree_leaf_box
        procedure :: leaf_u => robin_tree_leaf_u
    end type robin_tree

contains

    ! robin_tree%factor (robin_tree_module.f90:96-185)
    subroutine robin_tree_factor(T, L, ell_op, opts)
        class(robin_tree), intent(inout) :: T
        integer, intent(in) :: L
        type(elliptic_operator), intent(in) :: ell_op
        type(solve_opts), intent(in) :: opts
        type(rs_opts_t) :: co
        real(c_double) :: dom(6), h3(3)
        integer(c_int) :: pb(3), nl, m, ni, nb, rc
        complex(c_double_complex), allocatable :: kint(:,:)
        real(dp), allocatable :: pts(:,:)
        integer(c_int), allocatable :: fc(:)
        integer :: q, p, j

        call start_clock()
        T%nlvls = L
        T%ell_op = ell_op
        T%opts = opts
        if (c_associated(T%handle)) rc = rs_tree_destroy(T%handle)
        call rs_opts_default(co)
        co%precision = opts%precision
        co%n_refine = opts%n_refine
        rc = rs_tree_create(T%handle, co)
        call rs_check(rc, c_null_ptr)
        do j = 1, 3
            dom(2*j-1) = ell_op%domain(1, j)
            dom(2*j) = ell_op%domain(2, j)
        end do
        rc = rs_plan(T%handle, ell_op%d, L, dom, opts%h_tgt, opts%kb, T%bndry_params, nl, m, ni, nb)
        call rs_check(rc, T%handle)
        T%n_leaves = nl; T%m = m; T%n_int = ni; T%n_bnd = nb
        rc = rs_leaf_grid(T%handle, pb, h3)
        T%opts%h = h3
        write(*, '(A)') 'Domain partition completed.'
        allocate(kint(ni, nl), pts(3, m), fc(m))
        do q = 1, nl
            call leaf_points(T, q, pts, fc)
            do p = 1, ni
                kint(p, q) = ell_op%k(pts(:, nb + p))
            end do
        end do
        rc = rs_factor(T%handle, kint)
        call rs_check(rc, T%handle)
        write(*, '(A)') 'RTR computation completed.'
        T%factored = .true.
        write(*, '(A)') 'Factorization completed.'
        call read_clock(T%f_time)
    end subroutine robin_tree_factor

    subroutine leaf_points(T, q, pts, fc)
        class(robin_tree), intent(in) :: T
        integer, intent(in) :: q
        real(dp), intent(out) :: pts(:,:)
        integer(c_int), intent(out) :: fc(:)
        integer(c_int) :: rc
        rc = rs_leaf_points(T%handle, q - 1, pts, fc)
        call rs_check(rc, T%handle)
    end subroutine leaf_points

    ! robin_tree%solve (robin_tree_module.f90:198-233); repeatable
    subroutine robin_tree_solve(T, rhs)
        class(robin_tree), intent(inout) :: T
        type(elliptic_rhs), intent(in) :: rhs
        complex(c_double_complex), allocatable :: f(:,:), g(:,:)
        real(dp), allocatable :: pts(:,:)
        real(dp) :: box(2,3), x(3)
        integer(c_int), allocatable :: fc(:)
        integer(c_int) :: rc
        integer :: q, p, face, ax, j, nt, d
        logical :: phys

        call start_clock()
        if (.not. T%factored) then
            stop 'Error: Robin Tree must be factored first.'
        else if (T%ell_op%d /= rhs%d) then
            stop 'Error: Dimensions of Robin Tree and RHS do not match.'
        end if
        write(*, '(A)') 'Commencing forward substitution...'
        d = T%ell_op%d
        allocate(f(T%n_int, T%n_leaves), g(T%n_bnd, T%n_leaves), pts(3, T%m), fc(T%m))
        g = (0.0d0, 0.0d0)
        do q = 1, T%n_leaves
            call leaf_points(T, q, pts, fc)
            box = T%leaf_box(q)
            do p = 1, T%n_int
                f(p, q) = rhs%f(pts(:, T%n_bnd + p))
            end do
            ! boundary data on physical faces: fd_make_bndry_vec (fd_module.f90:611-659);
            ! g_face receives the tangential coordinates packed from x(1)
            do p = 1, T%n_bnd
                face = fc(p)
                ax = (face + 1) / 2
                if (mod(face, 2) == 1) then
                    phys = box(1, ax) == T%ell_op%domain(1, ax)
                else
                    phys = box(2, ax) == T%ell_op%domain(2, ax)
                end if
                if (.not. phys) cycle
                if (.not. associated(rhs%g(face)%g_face)) cycle
                x = 0.0d0
                nt = 0
                do j = 1, d
                    if (j /= ax) then
                        nt = nt + 1
                        x(nt) = pts(j, p)
                    end if
                end do
                g(p, q) = rhs%g(face)%g_face(x)
            end do
        end do
        write(*, '(A)') 'Commencing backward substitution...'
        rc = rs_solve(T%handle, 1, f, g, -1)
        call rs_check(rc, T%handle)
        write(*, '(A)') 'Elliptic solve completed.'
        call read_clock(T%s_time)
    end subroutine robin_tree_solve

    ! robin_tree%memsize: the reference's is empty (robin_tree_module.f90:243-250)
    subroutine robin_tree_memsize(T)
        class(robin_tree), intent(in) :: T
    end subroutine robin_tree_memsize

    ! robin_tree%inspect (robin_tree_module.f90:263-274)
    subroutine robin_tree_inspect(T)
        class(robin_tree), intent(in) :: T
        write(*, '(A,L1)') 'Factored: ', T%factored
        if (T%factored) then
            write(*, '(A,I0)') 'Number of levels: ', T%nlvls
        end if
    end subroutine robin_tree_inspect

    function robin_tree_leaf_box(T, q) result(box)
        class(robin_tree), intent(in) :: T
        integer, intent(in) :: q
        real(dp) :: box(2,3)
        real(c_double) :: b(6)
        integer(c_int) :: rc
        integer :: j
        rc = rs_leaf_box(T%handle, q - 1, b)
        call rs_check(rc, T%handle)
        do j = 1, 3
            box(1, j) = b(2*j-1)
            box(2, j) = b(2*j)
        end do
    end function robin_tree_leaf_box

    subroutine robin_tree_leaf_u(T, u_all)
        class(robin_tree), intent(in) :: T
        complex(c_double_complex), intent(out) :: u_all(:,:)
        integer(c_int) :: rc
        rc = rs_get_solution(T%handle, 0, u_all)
        call rs_check(rc, T%handle)
    end subroutine robin_tree_leaf_u

end module robin_tree_module


module robin_output_module
    use iso_c_binding
    use derived_type_module
    use robin_tree_module
    implicit none
    private
    integer, parameter :: dp = kind(1.0d0)

    public :: leaf_solution, get_solution, check_solution, write_variable

    type leaf_solution
        real(dp), dimension(1:2,1:3) :: box
        complex(dp), dimension(:), allocatable :: u
    end type leaf_solution

contains

    ! get_solution (robin_output_module.f90:48-75): leaves in DFS order
    subroutine get_solution(T, sol)
        type(robin_tree), intent(in) :: T
        type(leaf_solution), dimension(:), allocatable, intent(inout) :: sol
        complex(c_double_complex), allocatable :: u_all(:,:)
        integer :: q
        if (.not. T%factored) then
            stop 'Error in GET_SOLUTION: Robin Tree not factored.'
        end if
        allocate(u_all(T%m, T%n_leaves))
        call T%leaf_u(u_all)
        if (allocated(sol)) deallocate(sol)
        allocate(sol(T%n_leaves))
        do q = 1, T%n_leaves
            sol(q)%box = T%leaf_box(q)
            allocate(sol(q)%u(T%m))
            sol(q)%u = u_all(:, q)
        end do
        write(*, '(A, I0, A)') 'Solution vectors extracted on ', T%n_leaves, ' leaf nodes.'
    end subroutine get_solution

    ! check_solution (robin_output_module.f90:87-138): per-leaf error norms against true_sol
    ! evaluated at the leaf points (proc2vec, robin_output_module.f90:234-273)
    subroutine check_solution(d, leaf_sol, true_sol, opts, error_vec, norm, rel)
        integer, intent(in) :: d
        type(leaf_solution), dimension(:), intent(in) :: leaf_sol
        procedure(sol_proc) :: true_sol
        type(solve_opts), intent(in) :: opts
        real(dp), dimension(:), allocatable, intent(out) :: error_vec
        character(len=3), intent(in) :: norm
        logical, intent(in) :: rel
        complex(dp), allocatable :: ue(:)
        integer :: j
        allocate(error_vec(size(leaf_sol)))
        do j = 1, size(leaf_sol)
            call proc2vec(d, leaf_sol(j)%box, true_sol, ue, opts)
            if (.not. rel) then
                if (norm == '1') then
                    error_vec(j) = sum(abs(leaf_sol(j)%u - ue))
                else if (norm == '2') then
                    error_vec(j) = norm2(abs(leaf_sol(j)%u - ue))
                else
                    error_vec(j) = maxval(abs(leaf_sol(j)%u - ue))
                end if
            else
                if (norm == '1') then
                    error_vec(j) = sum(abs((leaf_sol(j)%u - ue) / ue))
                else if (norm == '2') then
                    error_vec(j) = norm2(abs((leaf_sol(j)%u - ue) / ue))
                else
                    error_vec(j) = maxval(abs((leaf_sol(j)%u - ue) / ue))
                end if
            end if
            deallocate(ue)
        end do
    end subroutine check_solution

    ! proc2vec: evaluate proc at the fgrid points of a box (grid from get_h, fd_module.f90:1060-1084)
    subroutine proc2vec(d, box, proc, vec_out, opts)
        integer, intent(in) :: d
        real(dp), dimension(1:2,1:3), intent(in) :: box
        procedure(sol_proc) :: proc
        complex(dp), dimension(:), allocatable, intent(out) :: vec_out
        type(solve_opts), intent(in) :: opts
        integer :: n(3), j, f, t, q, nb, ni, ax, rem(2), nr, p, g(3)
        real(dp) :: h(3), x(3)
        n = 1
        h = 0.0d0
        do j = 1, d
            n(j) = max(ceiling((box(2, j) - box(1, j)) / opts%h_tgt), 3)
            h(j) = (box(2, j) - box(1, j)) / n(j)
        end do
        ni = product(n(1:d))
        nb = 0
        do j = 1, d
            nb = nb + 2 * ni / n(j)
        end do
        allocate(vec_out(nb + ni))
        p = 0
        do f = 1, 2 * d
            ax = (f + 1) / 2
            nr = 0
            do j = 1, d
                if (j /= ax) then
                    nr = nr + 1
                    rem(nr) = j
                end if
            end do
            do t = 1, ni / n(ax)
                g = 1
                if (mod(f, 2) == 1) then
                    g(ax) = 0
                else
                    g(ax) = n(ax) + 1
                end if
                if (nr == 1) then
                    g(rem(1)) = t
                else
                    g(rem(1)) = mod(t - 1, n(rem(1))) + 1
                    g(rem(2)) = (t - 1) / n(rem(1)) + 1
                end if
                p = p + 1
                call pt(g)
                vec_out(p) = proc(x)
            end do
        end do
        do q = 1, ni
            g(1) = mod(q - 1, n(1)) + 1
            if (d == 2) then
                g(2) = (q - 1) / n(1) + 1
                g(3) = 1
            else
                g(2) = mod((q - 1) / n(1), n(2)) + 1
                g(3) = (q - 1) / (n(1) * n(2)) + 1
            end if
            p = p + 1
            call pt(g)
            vec_out(p) = proc(x)
        end do
    contains
        subroutine pt(gg)
            integer, intent(in) :: gg(3)
            integer :: jj
            do jj = 1, 3
                x(jj) = box(1, jj) + gg(jj) * h(jj) - h(jj) / 2
                if (jj <= d) then
                    if (gg(jj) == 0) x(jj) = box(1, jj)
                    if (gg(jj) == n(jj) + 1) x(jj) = box(2, jj)
                end if
            end do
        end subroutine pt
    end subroutine proc2vec

    ! write_variable: writes nothing in the reference either (robin_output_module.f90:157-221);
    ! kept with its argument list so that drivers calling it compile
    subroutine write_variable(T, unitval, filename, nodeid, variable, idx1, idx2)
        type(robin_tree), intent(in) :: T
        integer, intent(in) :: unitval
        character(len=*), intent(in) :: filename
        integer, intent(in) :: nodeid
        character(len=3), intent(in) :: variable
        integer, intent(in) :: idx1, idx2
        if (nodeid < 0 .or. nodeid > 2 * T%n_leaves - 2) then
            write(*, '(A)') 'Error in WRITE_VARIABLE: Argument nodeid out of bounds'
        end if
    end subroutine write_variable

end module robin_output_module
