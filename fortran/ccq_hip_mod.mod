﻿!mod$ v1 sum:d5f53c7c1d1eab6f
!need$ 0bde2ac47243ead2 i iso_c_binding
module ccq_hip_mod
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
type,bind(c)::ccq_hip_dims
integer(8)::froz
integer(8)::occ_a
integer(8)::occ_b
integer(8)::orbs
end type
type,bind(c)::ccq_hip_opts
integer(8)::level
integer(8)::rhf
real(8)::shift
integer(8)::diis_space
integer(8)::device
integer(8)::act_occ
integer(8)::act_unocc
end type
interface
function ccq_hip_create(dims,opts,ctx) bind(c,name="ccq_hip_create")
import::c_ptr
import::ccq_hip_dims
import::ccq_hip_opts
type(ccq_hip_dims),intent(in)::dims
type(ccq_hip_opts),intent(in)::opts
type(c_ptr),intent(out)::ctx
integer(4)::ccq_hip_create
end
end interface
interface
function ccq_hip_set_ints(ctx,f_a,f_b,v_aa,v_ab,v_bb) bind(c,name="ccq_hip_set_ints")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::f_a(1_8:*)
real(8),intent(in)::f_b(1_8:*)
real(8),intent(in)::v_aa(1_8:*)
real(8),intent(in)::v_ab(1_8:*)
real(8),intent(in)::v_bb(1_8:*)
integer(4)::ccq_hip_set_ints
end
end interface
interface
function ccq_hip_t_size(ctx) bind(c,name="ccq_hip_t_size")
import::c_ptr
type(c_ptr),value::ctx
integer(8)::ccq_hip_t_size
end
end interface
interface
function ccq_hip_set_t(ctx,t_vec,t_size) bind(c,name="ccq_hip_set_t")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::t_vec(1_8:*)
integer(8),value::t_size
integer(4)::ccq_hip_set_t
end
end interface
interface
function ccq_hip_get_t(ctx,t_vec,t_size) bind(c,name="ccq_hip_get_t")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::t_vec(1_8:*)
integer(8),value::t_size
integer(4)::ccq_hip_get_t
end
end interface
interface
function ccq_hip_t4_size(ctx) bind(c,name="ccq_hip_t4_size")
import::c_ptr
type(c_ptr),value::ctx
integer(8)::ccq_hip_t4_size
end
end interface
interface
function ccq_hip_set_t4(ctx,t4,n) bind(c,name="ccq_hip_set_t4")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::t4(1_8:*)
integer(8),value::n
integer(4)::ccq_hip_set_t4
end
end interface
interface
function ccq_hip_get_t4(ctx,t4,n) bind(c,name="ccq_hip_get_t4")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::t4(1_8:*)
integer(8),value::n
integer(4)::ccq_hip_get_t4
end
end interface
interface
function ccq_hip_update(ctx) bind(c,name="ccq_hip_update")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::ccq_hip_update
end
end interface
interface
function ccq_hip_energy(ctx,e) bind(c,name="ccq_hip_energy")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::e
integer(4)::ccq_hip_energy
end
end interface
interface
function ccq_hip_diis_record(ctx,iter) bind(c,name="ccq_hip_diis_record")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::iter
integer(4)::ccq_hip_diis_record
end
end interface
interface
function ccq_hip_residuum(ctx,iter,res) bind(c,name="ccq_hip_residuum")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::iter
real(8),intent(out)::res
integer(4)::ccq_hip_residuum
end
end interface
interface
function ccq_hip_diis_extrapolate(ctx) bind(c,name="ccq_hip_diis_extrapolate")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::ccq_hip_diis_extrapolate
end
end interface
interface
function ccq_hip_last_error(ctx) bind(c,name="ccq_hip_last_error")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::ccq_hip_last_error
end
end interface
interface
function ccq_hip_destroy(ctx) bind(c,name="ccq_hip_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::ccq_hip_destroy
end
end interface
contains
subroutine hip_check(ctx,rc,where)
type(c_ptr),intent(in)::ctx
integer(4),intent(in)::rc
character(*,1),intent(in)::where
end
end
