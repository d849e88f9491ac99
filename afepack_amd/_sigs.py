"""ctypes signatures of the C ABI declared in include/afepack_amd.h."""
import ctypes as C

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
U64 = C.c_uint64
F64 = C.c_double
PP = C.POINTER(C.c_void_p)
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)
PF64 = C.POINTER(C.c_double)


class Field(C.Structure):
    _fields_ = [("n_modes", I32), ("dim", I32), ("k", PF64), ("phi", PF64),
                ("w", PF64), ("scale", F64)]


class AmgParams(C.Structure):
    _fields_ = [("alpha", F64), ("nmpo", C.c_uint32), ("nmpsd", F64), ("lazy", I32),
                ("smooth", I32), ("build_fp32", I32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", I32), ("converged", I32), ("rel_res_l2", F64),
                ("rel_res_l1", F64), ("res_l1", F64), ("seconds", F64)]


class Request(C.Structure):
    _fields_ = [("mesh_key", I64), ("kind", I32), ("n", I32), ("jitter", F64), ("seed", U64),
                ("move_amp", F64), ("move_k", I32), ("acc", I32), ("op", I32), ("fkind", I32),
                ("coef", Field), ("rtol", F64), ("max_it", I32), ("keep_x", I32),
                ("n_vert", I64), ("coords", P), ("n_elem", I64), ("elem_vert", P), ("vert_mark", P)]


PField = C.POINTER(Field)

SIGS = {
    "afe_last_error": (C.c_char_p, []),
    "afe_version": (C.c_int, []),
    "afe_ctx_create": (C.c_int, [C.c_int, PP]),
    "afe_ctx_destroy": (C.c_int, [P]),
    "afe_ctx_synchronize": (C.c_int, [P]),
    "afe_ctx_device_bytes": (C.c_int, [P, PI64]),
    "afe_dev_alloc": (C.c_int, [P, I64, PP]),
    "afe_dev_free": (C.c_int, [P, P]),
    "afe_dev_upload": (C.c_int, [P, P, P, I64]),
    "afe_dev_download": (C.c_int, [P, P, P, I64]),
    "afe_dev_copy": (C.c_int, [P, P, P, I64]),
    "afe_dev_zero": (C.c_int, [P, P, I64]),
    "afe_mesh_square": (C.c_int, [I32, F64, U64, P, P, P]),
    "afe_mesh_cube": (C.c_int, [I32, F64, U64, P, P, P]),
    "afe_mesh_displace": (C.c_int, [I32, I64, P, P, F64, I32]),
    "afe_field_random": (C.c_int, [I32, I32, F64, U64, P, P, P, PF64]),
    "afe_space_create": (C.c_int, [P, I32, I64, P, I64, P, P, PP]),
    "afe_space_create_sides": (C.c_int, [P, I32, I64, P, I64, P, P, I64, P, P, PP]),
    "afe_space_destroy": (C.c_int, [P]),
    "afe_space_info": (C.c_int, [P, PI64, PI64, PI32]),
    "afe_space_dofs": (C.c_int, [P, P, P, P]),
    "afe_space_pattern": (C.c_int, [P, P, P]),
    "afe_space_update_coords": (C.c_int, [P, P]),
    "afe_matrix_create": (C.c_int, [P, PP]),
    "afe_matrix_create_csr": (C.c_int, [P, I64, P, P, P, PP]),
    "afe_matrix_create_pair": (C.c_int, [P, P, P]),
    "afe_matrix_info": (C.c_int, [P, P, P, P, P]),
    "afe_matrix_pattern": (C.c_int, [P, P, P]),
    "afe_matrix_destroy": (C.c_int, [P]),
    "afe_assemble": (C.c_int, [P, I32, I32, PField]),
    "afe_assemble_elements": (C.c_int, [P, P]),
    "afe_matrix_values": (C.c_int, [P, P]),
    "afe_matrix_set_values": (C.c_int, [P, P]),
    "afe_matrix_device_values": (C.c_void_p, [P]),
    "afe_matrix_vmult": (C.c_int, [P, P, P]),
    "afe_matrix_tvmult": (C.c_int, [P, P, P]),
    "afe_rhs": (C.c_int, [P, I32, I32, PField, P, P]),
    "afe_apply_dirichlet": (C.c_int, [P, P, P]),
    "afe_space_qpoints": (C.c_int, [P, I32, PI64, P, P]),
    "afe_rhs_qp": (C.c_int, [P, I32, P, P]),
    "afe_apply_dirichlet_values": (C.c_int, [P, P, I32, P, P, P]),
    "afe_apply_dirichlet_bc": (C.c_int, [P, P, I32, P, P, P, I32]),
    "afe_error_qp": (C.c_int, [P, P, I32, I32, P, PF64]),
    "afe_error_qp_noalloc": (C.c_int, [P, P, I32, I32, P, P]),
    "afe_error_l2": (C.c_int, [P, P, I32, PF64]),
    "afe_error_h1": (C.c_int, [P, P, I32, PF64]),
    "afe_amg_create": (C.c_int, [P, C.POINTER(AmgParams), PP]),
    "afe_amg_destroy": (C.c_int, [P]),
    "afe_amg_create_host": (C.c_int, [I64, P, P, P, C.POINTER(AmgParams), PP]),
    "afe_amg_lazy_reinit_host": (C.c_int, [P, P]),
    "afe_amg_coarse_inverse": (C.c_int, [P, P]),
    "afe_amg_lazy_reinit": (C.c_int, [P, P]),
    "afe_space_clone": (C.c_int, [P, P, PP]),
    "afe_amg_clone": (C.c_int, [P, P, PP]),
    "afe_amg_info": (C.c_int, [P, PI32, P, P, PI32]),
    "afe_amg_level_csr": (C.c_int, [P, I32, I32, PI64, PI64, PI64, P, P, P]),
    "afe_amg_level_colours": (C.c_int, [P, I32, P]),
    "afe_amg_solve": (C.c_int, [P, P, P, F64, I64, I32, F64, C.POINTER(SolveStats)]),
    "afe_amg_precondition": (C.c_int, [P, P, P, I32]),
    "afe_amg_pcg": (C.c_int, [P, P, P, F64, I32, I32, C.POINTER(SolveStats)]),
    "afe_amg_vcycle": (C.c_int, [P, P, P, I32]),
    "afe_amg_set_smooth": (C.c_int, [P, I32]),
    "afe_amg_set_coarse_exact": (C.c_int, [P, I32]),
    "afe_set_tuning": (C.c_int, [I32, I32, I64]),
    "afe_amg_kernel_timing": (C.c_int, [P, I32, I32, PF64, PF64]),
    "afe_assemble_kernel_timing": (C.c_int, [P, I32, I32, P, I32, PF64, PF64, PF64]),
    "afe_assemble_timing": (C.c_int, [P, I32, I32, P, I32, PF64]),
    "afe_copy_kernel_timing": (C.c_int, [P, I64, I32, PF64]),
    "afe_amg_set_tail_rows": (C.c_int, [P, I64]),
    "afe_amg_live_timing": (C.c_int, [P, I32, PF64, PI64]),
    "afe_amg_part_timing": (C.c_int, [P, I32, PF64, PF64, PI64]),
    "afe_server_create": (C.c_int, [I32, I64, I32, I32, I32, PP]),
    "afe_server_destroy": (C.c_int, [P]),
    "afe_server_submit": (C.c_int, [P, C.POINTER(Request), PI64]),
    "afe_server_poll": (C.c_int, [P, I64, PI32]),
    "afe_server_wait": (C.c_int, [P, I64, P, PI64, C.POINTER(SolveStats), PF64]),
    "afe_server_release": (C.c_int, [P, I64]),
    "afe_server_info": (C.c_int, [P, PI64, PI64, PI64, PI64]),
    "afe_server_timing": (C.c_int, [P, PF64]),
    "afe_server_timing_n": (C.c_int, [P, P, I32]),
    "afe_batch_create": (C.c_int, [P, I32, I32, PP]),
    "afe_batch_destroy": (C.c_int, [P]),
    "afe_batch_add": (C.c_int, [P, P, P, P, F64, I32, PI32]),
    "afe_batch_step": (C.c_int, [P, PI32]),
    "afe_batch_run": (C.c_int, [P]),
    "afe_batch_result": (C.c_int, [P, I32, PI32, P]),
    "afe_batch_remove": (C.c_int, [P, I32]),
    "afe_batch_set_tiles": (C.c_int, [P, I32, I32]),
    "afe_amg_device_info": (C.c_int, [P, I32, PI64, PI32, PI64, PI32, PI32]),
}


def declare(lib):
    for name, (res, args) in SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
