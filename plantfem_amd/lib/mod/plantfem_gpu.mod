﻿!mod$ v1 sum:321b6fadfa965882
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ f1de5abe9bfe2168 i iso_fortran_env
!need$ bb381bf46e508468 i __fortran_builtins
module plantfem_gpu
use,intrinsic::__fortran_builtins,only:__builtin_c_ptr
use,intrinsic::iso_fortran_env,only:int32
use,intrinsic::iso_fortran_env,only:int64
use,intrinsic::iso_fortran_env,only:real64
use,intrinsic::iso_fortran_env,only:error_unit
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
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::__builtin_c_ptr
private::int32
private::int64
private::real64
private::error_unit
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::pf_ok=0_4
integer(4),parameter::pf_breakdown_vv=1_4
integer(4),parameter::pf_breakdown_c2c3=2_4
integer(4),parameter::pf_not_converged=3_4
integer(4),parameter::pf_invalid=4_4
type,bind(c)::pf_solve_opts
integer(4)::itrmax
integer(4)::precision
real(8)::er0
real(8)::relative_er
real(8)::inner_tol
real(8)::target_relres
integer(4)::max_outer
integer(4)::spmv_kind
integer(4)::check_every
integer(4)::use_graph
integer(4)::debug
integer(4)::stop_mode
integer(4)::kernel_timing
integer(4)::jacobi
integer(4)::inner_max
integer(4)::reserved
end type
type,bind(c)::pf_stats
integer(4)::status
integer(4)::iterations
integer(4)::outer_iterations
integer(4)::spmv_kind_used
real(8)::rr
real(8)::rr1
real(8)::true_relres
real(8)::solve_ms
real(8)::setup_ms
real(8)::spmv_p_ms
real(8)::spmv_e_ms
real(8)::update_ms
real(8)::pupdate_ms
integer(8)::timed_iterations
end type
private::c_pf_solve_opts_default
interface
subroutine c_pf_solve_opts_default(o) bind(c,name="pf_solve_opts_default")
import::pf_solve_opts
type(pf_solve_opts),intent(inout)::o
end
end interface
private::c_pf_last_error
interface
function c_pf_last_error() bind(c,name="pf_last_error")
import::c_ptr
type(c_ptr)::c_pf_last_error
end
end interface
private::c_pf_version
interface
function c_pf_version() bind(c,name="pf_version")
integer(4)::c_pf_version
end
end interface
private::c_pf_ctx_create
interface
function c_pf_ctx_create(device,ctx) bind(c,name="pf_ctx_create")
import::c_ptr
integer(4),value::device
type(c_ptr),intent(out)::ctx
integer(4)::c_pf_ctx_create
end
end interface
private::c_pf_crs_graph
interface
function c_pf_crs_graph(ctx,nn,ne,nne,elemnod,dof,nnz,row_ptr,col_idx) bind(c,name="pf_crs_graph")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::nn
integer(8),value::ne
integer(4),value::nne
integer(4),intent(in)::elemnod(1_8:*)
integer(4),value::dof
integer(8),intent(out)::nnz
type(c_ptr),value::row_ptr
type(c_ptr),value::col_idx
integer(4)::c_pf_crs_graph
end
end interface
private::c_pf_hex8_stiffness
interface
function c_pf_hex8_stiffness(ctx,nn,ne,nodcoord,elemnod,e,nu,jmode,first,count,ke) bind(c,name="pf_hex8_stiffness")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::nn
integer(8),value::ne
real(8),intent(in)::nodcoord(1_8:*)
integer(4),intent(in)::elemnod(1_8:*)
real(8),intent(in)::e(1_8:*)
real(8),intent(in)::nu(1_8:*)
integer(4),value::jmode
integer(8),value::first
integer(8),value::count
real(8),intent(out)::ke(1_8:*)
integer(4)::c_pf_hex8_stiffness
end
end interface
private::c_pf_hex8_stiffness_crs
interface
function c_pf_hex8_stiffness_crs(ctx,nn,ne,nodcoord,elemnod,e,nu,jmode,row_ptr,col_idx,val) bind(c,name="pf_hex8_stiffness_crs")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::nn
integer(8),value::ne
real(8),intent(in)::nodcoord(1_8:*)
integer(4),intent(in)::elemnod(1_8:*)
real(8),intent(in)::e(1_8:*)
real(8),intent(in)::nu(1_8:*)
integer(4),value::jmode
integer(8),intent(in)::row_ptr(1_8:*)
integer(4),intent(in)::col_idx(1_8:*)
real(8),intent(out)::val(1_8:*)
integer(4)::c_pf_hex8_stiffness_crs
end
end interface
private::c_pf_spmv_crs
interface
function c_pf_spmv_crs(ctx,n,row_ptr,col_idx,val,x,y) bind(c,name="pf_spmv_crs")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::n
integer(8),intent(in)::row_ptr(1_8:*)
integer(4),intent(in)::col_idx(1_8:*)
real(8),intent(in)::val(1_8:*)
real(8),intent(in)::x(1_8:*)
real(8),intent(out)::y(1_8:*)
integer(4)::c_pf_spmv_crs
end
end interface
private::c_pf_bicgstab_crs
interface
function c_pf_bicgstab_crs(ctx,n,row_ptr,col_idx,val,rhs,x,opts,stats) bind(c,name="pf_bicgstab_crs")
import::c_ptr
import::pf_solve_opts
import::pf_stats
type(c_ptr),value::ctx
integer(8),value::n
integer(8),intent(in)::row_ptr(1_8:*)
integer(4),intent(in)::col_idx(1_8:*)
real(8),intent(in)::val(1_8:*)
real(8),intent(in)::rhs(1_8:*)
real(8),intent(inout)::x(1_8:*)
type(pf_solve_opts),intent(in)::opts
type(pf_stats),intent(out)::stats
integer(4)::c_pf_bicgstab_crs
end
end interface
private::c_pf_system_create
interface
function c_pf_system_create(ctx,sys) bind(c,name="pf_system_create")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr),intent(out)::sys
integer(4)::c_pf_system_create
end
end interface
private::c_pf_system_destroy
interface
function c_pf_system_destroy(sys) bind(c,name="pf_system_destroy")
import::c_ptr
type(c_ptr),value::sys
integer(4)::c_pf_system_destroy
end
end interface
private::c_pf_system_set_mesh
interface
function c_pf_system_set_mesh(sys,nn,ne,nne,nodcoord,elemnod) bind(c,name="pf_system_set_mesh")
import::c_ptr
type(c_ptr),value::sys
integer(8),value::nn
integer(8),value::ne
integer(4),value::nne
real(8),intent(in)::nodcoord(1_8:*)
integer(4),intent(in)::elemnod(1_8:*)
integer(4)::c_pf_system_set_mesh
end
end interface
private::c_pf_system_zero_matrix
interface
function c_pf_system_zero_matrix(sys,dof,nnz) bind(c,name="pf_system_zero_matrix")
import::c_ptr
type(c_ptr),value::sys
integer(4),value::dof
integer(8),intent(out)::nnz
integer(4)::c_pf_system_zero_matrix
end
end interface
private::c_pf_system_set_crs
interface
function c_pf_system_set_crs(sys,n,nnz,row_ptr,col_idx,val) bind(c,name="pf_system_set_crs")
import::c_ptr
type(c_ptr),value::sys
integer(8),value::n
integer(8),value::nnz
integer(8),intent(in)::row_ptr(1_8:*)
integer(4),intent(in)::col_idx(1_8:*)
real(8),intent(in)::val(1_8:*)
integer(4)::c_pf_system_set_crs
end
end interface
private::c_pf_system_get_crs
interface
function c_pf_system_get_crs(sys,row_ptr,col_idx,val) bind(c,name="pf_system_get_crs")
import::c_ptr
type(c_ptr),value::sys
type(c_ptr),value::row_ptr
type(c_ptr),value::col_idx
type(c_ptr),value::val
integer(4)::c_pf_system_get_crs
end
end interface
private::c_pf_system_stiffness
interface
function c_pf_system_stiffness(sys,e,nu,jmode) bind(c,name="pf_system_stiffness")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(in)::e(1_8:*)
real(8),intent(in)::nu(1_8:*)
integer(4),value::jmode
integer(4)::c_pf_system_stiffness
end
end interface
private::c_pf_system_add_element_matrices
interface
function c_pf_system_add_element_matrices(sys,count,ids,ke) bind(c,name="pf_system_add_element_matrices")
import::c_ptr
type(c_ptr),value::sys
integer(8),value::count
integer(4),intent(in)::ids(1_8:*)
real(8),intent(in)::ke(1_8:*)
integer(4)::c_pf_system_add_element_matrices
end
end interface
private::c_pf_system_add_element_vectors
interface
function c_pf_system_add_element_vectors(sys,count,ids,vec) bind(c,name="pf_system_add_element_vectors")
import::c_ptr
type(c_ptr),value::sys
integer(8),value::count
integer(4),intent(in)::ids(1_8:*)
real(8),intent(in)::vec(1_8:*)
integer(4)::c_pf_system_add_element_vectors
end
end interface
private::c_pf_system_set_rhs
interface
function c_pf_system_set_rhs(sys,rhs) bind(c,name="pf_system_set_rhs")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(in)::rhs(1_8:*)
integer(4)::c_pf_system_set_rhs
end
end interface
private::c_pf_system_get_rhs
interface
function c_pf_system_get_rhs(sys,rhs) bind(c,name="pf_system_get_rhs")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(out)::rhs(1_8:*)
integer(4)::c_pf_system_get_rhs
end
end interface
private::c_pf_system_add_traction
interface
function c_pf_system_add_traction(sys,box,pressure) bind(c,name="pf_system_add_traction")
import::c_ptr
type(c_ptr),value::sys
real(8),intent(in)::box(1_8:6_8)
real(8),value::pressure
integer(4)::c_pf_system_add_traction
end
end interface
private::c_pf_system_fix
interface
function c_pf_system_fix(sys,nids,ids,values,fixvalue) bind(c,name="pf_system_fix")
import::c_ptr
type(c_ptr),value::sys
integer(8),value::nids
integer(4),intent(in)::ids(1_8:*)
type(c_ptr),value::values
real(8),value::fixvalue
integer(4)::c_pf_system_fix
end
end interface
private::c_pf_system_solve
interface
function c_pf_system_solve(sys,opts,x0,x_out,stats) bind(c,name="pf_system_solve")
import::c_ptr
import::pf_solve_opts
import::pf_stats
type(c_ptr),value::sys
type(pf_solve_opts),intent(in)::opts
type(c_ptr),value::x0
type(c_ptr),value::x_out
type(pf_stats),intent(out)::stats
integer(4)::c_pf_system_solve
end
end interface
private::c_strlen
interface
function c_strlen(s) bind(c,name="strlen")
import::c_ptr
type(c_ptr),value::s
integer(8)::c_strlen
end
end interface
type::mesh_
real(8),allocatable::nodcoord(:,:)
integer(4),allocatable::elemnod(:,:)
end type
type::crs_
integer(4),allocatable::col_idx(:)
integer(8),allocatable::row_ptr(:)
real(8),allocatable::val(:)
end type
type::femdomain_
type(mesh_)::mesh
contains
procedure::create=>createfemdomain
procedure::nn=>nnfemdomain
procedure::ne=>nefemdomain
procedure::nne=>nnefemdomain
procedure::nd=>ndfemdomain
procedure::z_min=>z_minfemdomain
procedure::z_max=>z_maxfemdomain
procedure::x_min=>x_minfemdomain
procedure::x_max=>x_maxfemdomain
procedure::select=>selectfemdomain
procedure::zeromatrix=>zeromatrixfemdomain
procedure,pass::stiffnessmatrixelementfemdomain
procedure,pass::stiffnessmatrixcrsfemdomain
generic::stiffnessmatrix=>stiffnessmatrixelementfemdomain
generic::stiffnessmatrix=>stiffnessmatrixcrsfemdomain
end type
type::femsolver_
type(c_ptr)::sys=__builtin_c_ptr(__address=0_8)
type(femdomain_),pointer::domain=>NULL()
logical(4)::initialized=.false._4
logical(4)::debug=.false._4
integer(4)::itrmax=100000_4
real(8)::er0=1.00000001335143196001808973960578441619873046875e-10_8
real(8)::relative_er=1.00000001335143196001808973960578441619873046875e-10_8
integer(4)::dof=0_4
integer(8)::n=0_8
integer(8)::nnz=0_8
real(8),allocatable::crs_val(:)
real(8),allocatable::crs_rhs(:)
real(8),allocatable::crs_x(:)
integer(4),allocatable::crs_index_col(:)
integer(8),allocatable::crs_index_row(:)
type(pf_stats)::stats
integer(4),allocatable,private::pend_ids(:)
real(8),allocatable,private::pend_mat(:)
integer(8),private::npend=0_8
integer(8),private::mat_len=0_8
contains
procedure::init=>initfemsolver
procedure::setdomain=>setdomainfemsolver
procedure,pass::setcrsfemsolver
procedure,pass::setcrs_crsobjfemsolver
generic::setcrs=>setcrsfemsolver
generic::setcrs=>setcrs_crsobjfemsolver
procedure::setrhs=>setrhsfemsolver
procedure::getcrs=>getcrsfemsolver
procedure::setmatrix=>setmatrixfemsolver
procedure::setvector=>setvectorfemsolver
procedure::setstiffness=>setstiffnessfemsolver
procedure::addtraction=>addtractionfemsolver
procedure::fix=>fixfemsolver
procedure::solve=>solvefemsolver
procedure::flush=>flushfemsolver
procedure::remove=>removefemsolver
end type
intrinsic::null
private::null
intrinsic::dble
private::dble
type(c_ptr),private,save::g_ctx
private::last_error
private::check
private::ctx
private::default_opts
private::createfemdomain
private::nnfemdomain
private::nefemdomain
private::nnefemdomain
private::ndfemdomain
private::x_minfemdomain
private::x_maxfemdomain
private::z_minfemdomain
private::z_maxfemdomain
private::selectfemdomain
private::zeromatrixfemdomain
private::stiffnessmatrixelementfemdomain
private::stiffnessmatrixcrsfemdomain
private::initfemsolver
private::setdomainfemsolver
private::setcrsfemsolver
private::setcrs_crsobjfemsolver
private::setrhsfemsolver
private::flushfemsolver
private::setmatrixfemsolver
private::setvectorfemsolver
private::setstiffnessfemsolver
private::addtractionfemsolver
private::fixfemsolver
private::solvefemsolver
private::getcrsfemsolver
private::removefemsolver
contains
function last_error() result(msg)
character(:,1),allocatable::msg
end
subroutine check(code,where)
integer(4),intent(in)::code
character(*,1),intent(in)::where
end
function ctx() result(c)
type(c_ptr)::c
end
function pf_gpu_version()
integer(4)::pf_gpu_version
end
function default_opts() result(o)
type(pf_solve_opts)::o
end
subroutine createfemdomain(this,meshtype,x_num,y_num,z_num,x_len,y_len,z_len)
class(femdomain_),intent(inout)::this
character(*,1),intent(in)::meshtype
integer(4),intent(in),optional::x_num
integer(4),intent(in),optional::y_num
integer(4),intent(in),optional::z_num
real(8),intent(in),optional::x_len
real(8),intent(in),optional::y_len
real(8),intent(in),optional::z_len
end
function nnfemdomain(this)
class(femdomain_),intent(in)::this
integer(4)::nnfemdomain
end
function nefemdomain(this)
class(femdomain_),intent(in)::this
integer(4)::nefemdomain
end
function nnefemdomain(this)
class(femdomain_),intent(in)::this
integer(4)::nnefemdomain
end
function ndfemdomain(this)
class(femdomain_),intent(in)::this
integer(4)::ndfemdomain
end
function x_minfemdomain(this)
class(femdomain_),intent(in)::this
real(8)::x_minfemdomain
end
function x_maxfemdomain(this)
class(femdomain_),intent(in)::this
real(8)::x_maxfemdomain
end
function z_minfemdomain(this)
class(femdomain_),intent(in)::this
real(8)::z_minfemdomain
end
function z_maxfemdomain(this)
class(femdomain_),intent(in)::this
real(8)::z_maxfemdomain
end
function selectfemdomain(this,x_min,x_max,y_min,y_max,z_min,z_max) result(nodelist)
class(femdomain_),intent(in)::this
real(8),intent(in),optional::x_min
real(8),intent(in),optional::x_max
real(8),intent(in),optional::y_min
real(8),intent(in),optional::y_max
real(8),intent(in),optional::z_min
real(8),intent(in),optional::z_max
integer(4),allocatable::nodelist(:)
end
function zeromatrixfemdomain(this,dof) result(zeromatrix)
class(femdomain_),intent(in)::this
integer(4),intent(in)::dof
type(crs_)::zeromatrix
end
function stiffnessmatrixelementfemdomain(this,elementid,e,v) result(stiffnessmatrix)
class(femdomain_),intent(in)::this
integer(4),intent(in)::elementid
real(8),intent(in)::e
real(8),intent(in)::v
real(8),allocatable::stiffnessmatrix(:,:)
end
function stiffnessmatrixcrsfemdomain(this,youngmodulus,poissonratio) result(stiffnessmatrix)
class(femdomain_),intent(in)::this
real(8),intent(in)::youngmodulus(:)
real(8),intent(in)::poissonratio(:)
type(crs_)::stiffnessmatrix
end
subroutine initfemsolver(this,numdomain)
class(femsolver_),intent(inout)::this
integer(4),intent(in),optional::numdomain
end
subroutine setdomainfemsolver(this,femdomain,domainid)
class(femsolver_),intent(inout)::this
type(femdomain_),intent(in),target::femdomain
integer(4),intent(in),optional::domainid
end
subroutine setcrsfemsolver(this,dof)
class(femsolver_),intent(inout)::this
integer(4),intent(in)::dof
end
subroutine setcrs_crsobjfemsolver(this,crs)
class(femsolver_),intent(inout)::this
type(crs_),intent(in)::crs
end
subroutine setrhsfemsolver(this,rhs)
class(femsolver_),intent(inout)::this
real(8),intent(in)::rhs(:)
end
subroutine flushfemsolver(this)
class(femsolver_),intent(inout)::this
end
subroutine setmatrixfemsolver(this,domainid,elementid,dof,matrix)
class(femsolver_),intent(inout)::this
integer(4),intent(in)::domainid
integer(4),intent(in)::elementid
integer(4),intent(in)::dof
real(8),intent(in)::matrix(:,:)
end
subroutine setvectorfemsolver(this,domainid,elementid,dof,vector)
class(femsolver_),intent(inout)::this
integer(4),intent(in)::domainid
integer(4),intent(in)::elementid
integer(4),intent(in)::dof
real(8),intent(in)::vector(:)
end
subroutine setstiffnessfemsolver(this,youngmodulus,poissonratio)
class(femsolver_),intent(inout)::this
real(8),intent(in)::youngmodulus(:)
real(8),intent(in)::poissonratio(:)
end
subroutine addtractionfemsolver(this,box,surfacepressure)
class(femsolver_),intent(inout)::this
real(8),intent(in)::box(1_8:6_8)
real(8),intent(in)::surfacepressure
end
subroutine fixfemsolver(this,domainid,ids,fixvalue,fixvalues)
class(femsolver_),intent(inout)::this
integer(4),intent(in),optional::domainid
integer(4),intent(in)::ids(:)
real(8),intent(in),optional::fixvalue
real(8),intent(in),optional,target::fixvalues(:)
end
function solvefemsolver(this,x0,precision) result(x)
class(femsolver_),intent(inout)::this
real(8),intent(in),optional,target::x0(:)
integer(4),intent(in),optional::precision
real(8),allocatable::x(:)
end
subroutine getcrsfemsolver(this)
class(femsolver_),intent(inout),target::this
end
subroutine removefemsolver(this)
class(femsolver_),intent(inout)::this
end
subroutine pf_gpu_bicgstab(row_ptr,col_idx,val,rhs,x)
integer(8),intent(in)::row_ptr(:)
integer(4),intent(in)::col_idx(:)
real(8),intent(in)::val(:)
real(8),intent(in)::rhs(:)
real(8),intent(inout)::x(:)
end
function pf_gpu_bicgstab_func(row_ptr,col_idx,val,rhs,x0) result(x)
integer(8),intent(in)::row_ptr(:)
integer(4),intent(in)::col_idx(:)
real(8),intent(in)::val(:)
real(8),intent(in)::rhs(:)
real(8),intent(in)::x0(:)
real(8),allocatable::x(:)
end
subroutine pf_gpu_spmv(row_ptr,col_idx,val,old_vector,new_vector)
integer(8),intent(in)::row_ptr(:)
integer(4),intent(in)::col_idx(:)
real(8),intent(in)::val(:)
real(8),intent(in)::old_vector(:)
real(8),allocatable,intent(inout)::new_vector(:)
end
end
