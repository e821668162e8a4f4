! ISO_C_BINDING interface to the MI355X amplitude engine (include/ccq_hip.h).
!
! ccq binds its CC sweep and energy through update_ptr / calculate_energy_ptr
! (src/proc_pointers.f90:8-32) and calls write_vecs / residuum / calc_diis
! from jacobi_iter (src/solve_cc.f90:175-269).  These interfaces let a ccq
! build call the resident GPU engine at exactly those points.  ccq is built
! with -fdefault-integer-8, so every integer here is c_int64_t.
module ccq_hip_mod

    use, intrinsic :: iso_c_binding

    implicit none

    type, bind(c) :: ccq_hip_dims
        integer(c_int64_t) :: froz, occ_a, occ_b, orbs
    end type ccq_hip_dims

    type, bind(c) :: ccq_hip_opts
        integer(c_int64_t) :: level, rhf
        real(c_double) :: shift
        integer(c_int64_t) :: diis_space, device
        integer(c_int64_t) :: act_occ, act_unocc
    end type ccq_hip_opts

    interface

        integer(c_int) function ccq_hip_create(dims, opts, ctx) bind(c, name='ccq_hip_create')
            import :: c_int, c_ptr, ccq_hip_dims, ccq_hip_opts
            type(ccq_hip_dims), intent(in) :: dims
            type(ccq_hip_opts), intent(in) :: opts
            type(c_ptr), intent(out) :: ctx
        end function ccq_hip_create

        integer(c_int) function ccq_hip_set_ints(ctx, f_a, f_b, v_aa, v_ab, v_bb) &
                bind(c, name='ccq_hip_set_ints')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: f_a(*), f_b(*), v_aa(*), v_ab(*), v_bb(*)
        end function ccq_hip_set_ints

        integer(c_int) function ccq_hip_set_irreps(ctx, irreps, n) bind(c, name='ccq_hip_set_irreps')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(in) :: irreps(*)
            integer(c_int64_t), value :: n
        end function ccq_hip_set_irreps

        integer(c_int64_t) function ccq_hip_t_size(ctx) bind(c, name='ccq_hip_t_size')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_t_size

        integer(c_int) function ccq_hip_set_t(ctx, t_vec, t_size) bind(c, name='ccq_hip_set_t')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: t_vec(*)
            integer(c_int64_t), value :: t_size
        end function ccq_hip_set_t

        integer(c_int) function ccq_hip_get_t(ctx, t_vec, t_size) bind(c, name='ccq_hip_get_t')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: t_vec(*)
            integer(c_int64_t), value :: t_size
        end function ccq_hip_get_t

        integer(c_int64_t) function ccq_hip_t4_size(ctx) bind(c, name='ccq_hip_t4_size')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_t4_size

        integer(c_int) function ccq_hip_set_t4(ctx, t4, n) bind(c, name='ccq_hip_set_t4')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: t4(*)
            integer(c_int64_t), value :: n
        end function ccq_hip_set_t4

        integer(c_int) function ccq_hip_get_t4(ctx, t4, n) bind(c, name='ccq_hip_get_t4')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: t4(*)
            integer(c_int64_t), value :: n
        end function ccq_hip_get_t4

        integer(c_int) function ccq_hip_update(ctx) bind(c, name='ccq_hip_update')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_update

        integer(c_int) function ccq_hip_energy(ctx, e) bind(c, name='ccq_hip_energy')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: e
        end function ccq_hip_energy

        integer(c_int) function ccq_hip_diis_record(ctx, iter) bind(c, name='ccq_hip_diis_record')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
        end function ccq_hip_diis_record

        integer(c_int) function ccq_hip_residuum(ctx, iter, res) bind(c, name='ccq_hip_residuum')
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
            real(c_double), intent(out) :: res
        end function ccq_hip_residuum

        integer(c_int) function ccq_hip_diis_extrapolate(ctx) bind(c, name='ccq_hip_diis_extrapolate')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_diis_extrapolate

        ! CC(t;3) tail: hbar2 (src/hbar.f90:10-720), solve_lcc / update_l2
        ! (src/solve_cc.f90:101-173, src/update_lcc.f90:10-280) and crcc23
        ! (src/mm_correct.f90:11-343)
        integer(c_int) function ccq_hip_hbar(ctx) bind(c, name='ccq_hip_hbar')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_hbar

        integer(c_int) function ccq_hip_left_init(ctx, e_cor) bind(c, name='ccq_hip_left_init')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: e_cor
        end function ccq_hip_left_init

        integer(c_int) function ccq_hip_left_update(ctx) bind(c, name='ccq_hip_left_update')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_left_update

        integer(c_int) function ccq_hip_left_energy(ctx, e) bind(c, name='ccq_hip_left_energy')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: e
        end function ccq_hip_left_energy

        integer(c_int) function ccq_hip_left_mode(ctx, on) bind(c, name='ccq_hip_left_mode')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: on
        end function ccq_hip_left_mode

        integer(c_int64_t) function ccq_hip_l_size(ctx) bind(c, name='ccq_hip_l_size')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_l_size

        integer(c_int) function ccq_hip_set_l(ctx, l_vec, l_size) bind(c, name='ccq_hip_set_l')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: l_vec(*)
            integer(c_int64_t), value :: l_size
        end function ccq_hip_set_l

        integer(c_int) function ccq_hip_get_l(ctx, l_vec, l_size) bind(c, name='ccq_hip_get_l')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: l_vec(*)
            integer(c_int64_t), value :: l_size
        end function ccq_hip_get_l

        integer(c_int) function ccq_hip_cct3(ctx, e23) bind(c, name='ccq_hip_cct3')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: e23(4)
        end function ccq_hip_cct3

        type(c_ptr) function ccq_hip_last_error(ctx) bind(c, name='ccq_hip_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_last_error

        integer(c_int) function ccq_hip_destroy(ctx) bind(c, name='ccq_hip_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function ccq_hip_destroy

    end interface

contains

    subroutine hip_check(ctx, rc, where)
        ! Map a non-zero return code onto ccq's stop_all convention
        ! (src/errors.f90:11-27): message on stderr and exit status 1.
        type(c_ptr), intent(in) :: ctx
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc == 0) return
        call c_f_pointer(ccq_hip_last_error(ctx), msg, [1024])
        n = 0
        do while (n < 1024)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        end do
        write(0, '(a)') 'ERROR in '//where//': '//transfer(msg(1:n), repeat(' ', n))
        call exit(1)
    end subroutine hip_check

end module ccq_hip_mod
