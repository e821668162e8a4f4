! Minimal ccq-compatible driver for the GPU amplitude engine.
!
! Reads a ccq keyword file (src/parser.f90 semantics for the keywords used by
! the CC iteration), the onebody/twobody integral files (src/integrals.f90:
! 11-91), builds f_a, f_b, v_aa, v_ab, v_bb and E_HF exactly like ccq
! (src/integrals.f90:212-364, src/energy.f90:49-101), then runs ccq's
! jacobi_iter (src/solve_cc.f90:175-269) with every amplitude operation on
! the GPU through ccq_hip_mod, printing ccq's iteration and summary lines
! (src/printing.F90:303-349, 260-284).  calc_type=CCT3 adds the CC(t;3) tail
! of calc_driver (src/calc_driver.f90:53-72): H-bar, the left-CCSD
! iteration through the same jacobi_iter, crcc23 and print_cct3.
program ccq_drv

    use, intrinsic :: iso_c_binding
    use ccq_hip_mod

    implicit none

    integer, parameter :: dp = c_double
    character(len=512) :: inp, line, key, val, onebody, twobody, calc_type, sym_file
    integer(c_int64_t) :: froz_spin, nel, nvirt, mult, froz, occ_a, occ_b, orbs
    integer(c_int64_t) :: diis_space, max_iter, level, rhf_flag, act_occ, act_unocc
    real(dp) :: tol, shift
    real(dp), allocatable :: e1int(:,:), e2int(:,:,:,:), f_a(:,:), f_b(:,:)
    real(dp), allocatable :: v_aa(:,:,:,:), v_ab(:,:,:,:)
    real(dp) :: en_repul, en_ref, e_cor, e_left, e23(4), t0
    integer(c_int64_t) :: i, j, k, a, b, ios, unit_in, itol, eq, nacc
    logical :: rhf, converged, cct3, sym_exists
    integer(c_int64_t), allocatable :: orb_syms(:)
    integer(c_int64_t) :: point_group, unit_sym
    real :: accv(5)
    type(ccq_hip_dims) :: dims
    type(ccq_hip_opts) :: opts
    type(c_ptr) :: ctx

    call get_command_argument(1, inp)
    if (len_trim(inp) == 0) then
        write(0, '(a)') 'usage: ccq_drv input.inp'
        call exit(1)
    end if

    ! defaults (src/parser.f90:779-844)
    froz_spin = 0; nel = -1; nvirt = -1; mult = 1; act_occ = -1; act_unocc = -1
    cct3 = .false.
    tol = 1.0e-4_dp; diis_space = 5; max_iter = 60; shift = 0.0_dp; rhf = .false.
    onebody = 'onebody.inp'; twobody = 'twobody.inp'; calc_type = 'CCSD'; sym_file = ''

    open(newunit=unit_in, file=trim(inp), status='old', action='read')
    do
        read(unit_in, '(a)', iostat=ios) line
        if (ios /= 0) exit
        if (index(line, '#') + index(line, '!') /= 0) cycle
        if (len_trim(line) == 0) cycle
        eq = index(line, '=')
        if (eq == 0) cycle
        key = adjustl(line(1:eq - 1))
        val = adjustl(line(eq + 1:))
        select case (trim(key))
        case ('core', 'nfroz', 'frozen', 'froz'); read(val, *) froz_spin
        case ('nel', 'electrons'); read(val, *) nel
        case ('nvir', 'nvirt', 'virtuals'); read(val, *) nvirt
        case ('multiplicity', 'mult'); read(val, *) mult
        case ('rhf'); read(val, *) rhf
        case ('shift'); read(val, *) shift
        case ('max_iterations', 'max_iter'); read(val, *) max_iter
        case ('diis_space', 'diis'); read(val, *) diis_space
        case ('tolerance', 'tol', 'itol')
            read(val, '(i10)', iostat=ios) itol
            if (ios /= 0) then
                read(val, *) tol
            else
                tol = 10.0_dp ** itol
            end if
        case ('symfile', 'sym_file'); sym_file = val
        case ('onebody'); onebody = val
        case ('twobody'); twobody = val
        case ('calc_type'); calc_type = val
        case ('act_occ'); read(val, *) act_occ
        case ('act_unocc'); read(val, *) act_unocc
        case ('t2t2_t2', 't3_t2', 't1t3_t2', 't2t2_t3', 't2t3_t3')
            ! ACC diagram scale factors (src/parser.f90:529-543, read as
            ! REAL*4 lists of 5, 2, 4, 3, 5 values): this engine computes
            ! plain CC, the default 1.0 (src/parser.f90:838-842), only
            select case (trim(key))
            case ('t3_t2'); nacc = 2
            case ('t1t3_t2'); nacc = 4
            case ('t2t2_t3'); nacc = 3
            case default; nacc = 5
            end select
            accv = 1.0
            read(val, *, iostat=ios) accv(1:nacc)
            if (ios /= 0 .or. any(accv(1:nacc) /= 1.0)) then
                write(0, '(a)') 'ccq_drv: ACC scale factors '//trim(key)//' = '//trim(val)// &
                    ' are not supported (only the default 1.0)'
                call exit(1)
            end if
        end select
    end do
    close(unit_in)

    ! src/parser.f90:476-504
    froz = froz_spin / 2
    occ_a = nel / 2
    occ_b = occ_a - (mult - 1) / 2
    orbs = (nel + nvirt) / 2
    select case (trim(calc_type))
    case ('CCSD'); level = 2
    case ('CCSDT'); level = 3
    case ('CCSDTQ'); level = 4
    case ('CCSDt', 'CCT3', 'CCt3', 'CC(t;3)')
        level = 5
        cct3 = trim(calc_type) /= 'CCSDt'
        if (cct3) then
            level = 6
            calc_type = 'CC(t;3)'
        end if
        ! src/parser.f90:251-254
        if (act_occ == -1 .and. act_unocc == -1) then
            write(0, '(a)') 'CONFIGURATION ERROR: Active space methods require an active space'
            call exit(1)
        end if
    case default
        write(0, '(a)') 'ccq_drv: calc_type '//trim(calc_type)//' is not supported'
        call exit(1)
    end select

    allocate(e1int(orbs, orbs), e2int(orbs, orbs, orbs, orbs))
    e1int = 0; e2int = 0; en_repul = 0
    open(newunit=unit_in, file=trim(onebody), status='old', action='read')
    do i = 1, orbs
        do j = 1, i
            read(unit_in, *) e1int(i, j)
            e1int(j, i) = e1int(i, j)
        end do
    end do
    close(unit_in)
    open(newunit=unit_in, file=trim(twobody), status='old', action='read')
    do
        read(unit_in, *, iostat=ios) i, a, j, b, t0
        if (ios /= 0) exit
        if (i + a + j + b == 0) then
            en_repul = t0
            exit
        end if
        e2int(i, j, a, b) = t0
    end do
    close(unit_in)

    allocate(v_aa(froz+1:orbs, froz+1:orbs, froz+1:orbs, froz+1:orbs))
    allocate(v_ab(froz+1:orbs, froz+1:orbs, froz+1:orbs, froz+1:orbs))
    do b = froz + 1, orbs
        do a = froz + 1, orbs
            do j = froz + 1, orbs
                do i = froz + 1, orbs
                    v_ab(i, j, a, b) = e2int(i, j, a, b)
                    v_aa(i, j, a, b) = e2int(i, j, a, b) - e2int(i, j, b, a)
                end do
            end do
        end do
    end do
    allocate(f_a(orbs, orbs), f_b(orbs, orbs))
    f_a = e1int; f_b = e1int
    do i = 1, orbs
        do j = 1, i
            do k = 1, occ_b
                f_a(i, j) = f_a(i, j) + e2int(i, k, j, k)
            end do
            do k = 1, occ_a
                f_a(i, j) = f_a(i, j) + (e2int(i, k, j, k) - e2int(i, k, k, j))
                f_b(i, j) = f_b(i, j) + e2int(i, k, j, k)
            end do
            do k = 1, occ_b
                f_b(i, j) = f_b(i, j) + (e2int(i, k, j, k) - e2int(i, k, k, j))
            end do
            f_a(j, i) = f_a(i, j)
            f_b(j, i) = f_b(i, j)
        end do
    end do
    en_ref = 0
    do i = 1, occ_a
        en_ref = en_ref + e1int(i, i)
    end do
    do i = 1, occ_b
        en_ref = en_ref + e1int(i, i)
    end do
    do i = 1, occ_a
        do j = 1, occ_b
            en_ref = en_ref + e2int(i, j, i, j)
        end do
    end do
    do i = 1, occ_a
        do j = 1, occ_a
            en_ref = en_ref + 0.5_dp * (e2int(i, j, i, j) - e2int(i, j, j, i))
        end do
    end do
    do i = 1, occ_b
        do j = 1, occ_b
            en_ref = en_ref + 0.5_dp * (e2int(i, j, i, j) - e2int(i, j, j, i))
        end do
    end do
    en_ref = en_ref + en_repul

    dims = ccq_hip_dims(froz, occ_a, occ_b, orbs)
    rhf_flag = merge(1_c_int64_t, 0_c_int64_t, rhf)
    opts = ccq_hip_opts(level, rhf_flag, shift, diis_space, 0_c_int64_t, act_occ, act_unocc)
    if (ccq_hip_create(dims, opts, ctx) /= 0) then
        write(0, '(a)') 'ERROR in ccq_hip_create'
        call exit(1)
    end if
    ! orbital irreps (read_sym, src/symmetry.f90:15-119: the point group, then
    ! one irrep per orbital; a missing file is C1): the engine runs every
    ! product in its irrep blocks
    if (len_trim(sym_file) > 0) then
        inquire(file=trim(sym_file), exist=sym_exists)
        if (sym_exists) then
            allocate(orb_syms(orbs))
            open(newunit=unit_sym, file=trim(sym_file), status='old', action='read')
            read(unit_sym, *) point_group
            if (point_group < 1 .or. point_group > 4) then
                write(0, '(a)') 'ERROR in read_sym: Point-group symmetry not supported'
                call exit(1)
            end if
            do i = 1, orbs
                read(unit_sym, *) orb_syms(i)
            end do
            close(unit_sym)
            call hip_check(ctx, ccq_hip_set_irreps(ctx, orb_syms, orbs), 'set_irreps')
        end if
    end if
    call hip_check(ctx, ccq_hip_set_ints(ctx, f_a, f_b, v_aa, v_ab, v_aa), 'set_ints')

    write(*, '(a)') 'Coupled cluster'
    write(*, '(a)') '---------------'
    call jacobi_iter(.false., e_cor, converged)

    if (cct3) then
        ! hbar2, then solve_lcc with update_l2 / calculate_left_energy
        call hip_check(ctx, ccq_hip_hbar(ctx), 'hbar2')
        write(*, '(a)') 'Left Coupled cluster'
        write(*, '(a)') '--------------------'
        call hip_check(ctx, ccq_hip_left_init(ctx, e_cor), 'solve_lcc')
        call hip_check(ctx, ccq_hip_left_mode(ctx, 1_c_int64_t), 'solve_lcc')
        call jacobi_iter(.true., e_left, converged)
        call hip_check(ctx, ccq_hip_left_mode(ctx, 0_c_int64_t), 'solve_lcc')
        ! crcc23 and print_cct3 (src/printing.F90:234-258)
        call hip_check(ctx, ccq_hip_cct3(ctx, e23), 'crcc23')
        write(*, '(/a20,2a18)') 'Method', 'Correlation (Eh)', 'Total Energy (Eh)'
        write(*, '(10x,46("-"))')
        write(*, '(a20,2f18.12)') 'CCSDt', e_cor, en_ref + e_cor
        write(*, '(a20,2f18.12)') 'CC(t;3),A', e_cor + e23(1), en_ref + e_cor + e23(1)
        write(*, '(a20,2f18.12)') 'CC(t;3),B', e_cor + e23(2), en_ref + e_cor + e23(2)
        write(*, '(a20,2f18.12)') 'CC(t;3),C', e_cor + e23(3), en_ref + e_cor + e23(3)
        write(*, '(a20,2f18.12)') 'CC(t;3)', e_cor + e23(4), en_ref + e_cor + e23(4)
        e_cor = e_cor + e23(4)
    end if

    write(*, '(/a)') 'Calculation Summary (Eh)'
    write(*, '(a)') '------------------------'
    write(*, '(a20,f18.12)') 'Reference', en_ref
    write(*, '(a20,f18.12)') 'Correlation', e_cor
    write(*, '(a20,f18.12)') trim(calc_type), en_ref + e_cor
    i = ccq_hip_destroy(ctx)

contains

    ! jacobi_iter (src/solve_cc.f90:175-269) over the engine: the T equations,
    ! or with left the left-CC equations (update_l2 / calculate_left_energy)
    subroutine jacobi_iter(left, e_out, converged)
        logical, intent(in) :: left
        real(dp), intent(out) :: e_out
        logical, intent(out) :: converged
        real(dp) :: e_new, e_old, diff, res, conv(3), t0, t1, nsec
        integer(c_int64_t) :: it
        integer :: nmin
        logical :: stop_file
        write(*, '(/2x,a4,3(a15),a16)') 'It.', 'E (Corr)', 'dE', 'Residuum', 'Wall Time'
        write(*, '(2x,65("-"))')
        conv = 1.0_dp
        e_old = 0.0_dp
        converged = .false.
        t0 = wall_time()
        do it = 1, max_iter
            if (left) then
                call hip_check(ctx, ccq_hip_left_update(ctx), 'update_l2')
                call hip_check(ctx, ccq_hip_left_energy(ctx, e_new), 'calculate_left_energy')
            else
                call hip_check(ctx, ccq_hip_update(ctx), 'update')
                call hip_check(ctx, ccq_hip_energy(ctx, e_new), 'energy')
            end if
            diff = e_old - e_new
            conv(1) = conv(2); conv(2) = conv(3); conv(3) = diff
            e_old = e_new
            call hip_check(ctx, ccq_hip_diis_record(ctx, it), 'diis_record')
            call hip_check(ctx, ccq_hip_residuum(ctx, it, res), 'residuum')
            if (mod(it, diis_space + 1) == 0) then
                write(*, '(a)') '      DIIS cycle'
                call hip_check(ctx, ccq_hip_diis_extrapolate(ctx), 'calc_diis')
            end if
            ! wall time of the iteration split into min:sec (print_iteration,
            ! src/printing.F90:336-345; get_wall_time, src/utils/utils.F90:676)
            t1 = wall_time()
            nsec = t1 - t0
            nmin = int(nsec) / 60
            nsec = nsec - real(nmin, dp) * 60.0_dp
            write(*, '(2x,i4,3(f15.10),i5,'' min'',f5.1,'' s'')') it, e_new, diff, res, nmin, nsec
            t0 = t1
            ! external stop request: a file cc.comm is consumed and ends the
            ! loop (check_sig, src/solve_cc.f90:246-249,271-291)
            inquire(file='cc.comm', exist=stop_file)
            if (stop_file) then
                open(77, file='cc.comm', status='old')
                close(77, status='delete')
                write(*, '(a)') 'Signaled to end'
                exit
            end if
            if (all(abs(conv) < tol)) then
                converged = .true.
                exit
            end if
            if (e_new /= e_new) then
                write(*, '(a)') 'FAILED TO CONVERGE'
                exit
            end if
            if (it == max_iter) write(*, '(a)') 'FAILED TO CONVERGE IN', max_iter, &
                ' ITERATIONS. KEEPING LAST T VECTOR.'
        end do
        e_out = e_old
    end subroutine jacobi_iter

    function wall_time() result(t)
        real(dp) :: t
        integer(c_int64_t) :: t_int, count_rate, count_max
        call system_clock(t_int, count_rate, count_max)
        t = real(t_int, dp) / real(count_rate, dp)
    end function wall_time

end program ccq_drv
