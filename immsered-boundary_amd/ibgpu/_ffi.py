"""ctypes mirror of include/ibgpu.h and csrc/host/ib_mesh.h.

Loads the in-tree shared libraries built by ``immsered-boundary_amd/build.py``:
``libibhost.so`` (host mesh builder), ``libibgpu.so`` (HIP engine).  Nothing
here falls back to another implementation: a missing library raises.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)               # immsered-boundary_amd/
LIB_DIR = os.path.join(ROOT, "lib")

IBG_MAX_LEVELS = 12
IBG_MAX_CASES = 64
IBG_NFV = 7
IBM_MAX_BODIES = 8

i32p = C.POINTER(C.c_int32)
i8p = C.POINTER(C.c_int8)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
IBG_CLOSURE_NPARAM = 4801
IBG_CLOSURE_NFEAT = 8
IBG_CLOSURE_NOALLOC_MAX_ROWS = 65536


class MeshSoA(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("ncases", C.c_int32), ("nlevels", C.c_int32),
        ("ncells", C.c_int32),
        ("blk_start", i32p), ("blk_nnormal", i32p), ("blk_h", f64p),
        ("blk_twod", i32p),
        ("key", i32p), ("kind", i8p), ("nb", i32p), ("cell_faces", i32p),
        ("signdis", f64p), ("ibidx", i32p),
        ("nfaces", C.c_int32), ("blk_face_start", i32p), ("face_cells", i32p),
        ("face_dir", i8p),
        ("ndmg", C.c_int32), ("dmg_cell", i32p), ("dmg_ref", i32p),
        ("dmg_type", i32p),
        ("nres", C.c_int32), ("res_ci", i32p), ("res_fi", i32p), ("res_blk", i32p),
        ("npro", C.c_int32), ("pro_ci", i32p), ("pro_fi", i32p), ("pro_blk", i32p),
        ("nib", C.c_int32), ("ib_cell", i32p), ("ib_intp", i32p),
        ("ib_coef", f64p), ("ib_hgdis", f64p), ("ib_nv", f64p),
        ("ib_attach", f64p), ("ib_body", i32p),
        ("npatch", C.c_int32), ("pat_intp", i32p), ("pat_coef", f64p),
        ("pat_hgdis", f64p), ("pat_nv", f64p), ("pat_pc", f64p),
        ("pat_area", f64p), ("pat_case", i32p),
        ("nbody", C.c_int32), ("body_case", i32p), ("body_center", f64p),
        ("dmg_nbr", i32p), ("dmg_nmv", f64p), ("dist_dm", f64p),
        ("nmdy", C.c_int32), ("mdy", i32p),
    ]


class Physics(C.Structure):
    _fields_ = [
        ("gama", C.c_double), ("Ma", C.c_double), ("Re", C.c_double),
        ("Pr", C.c_double), ("Rg", C.c_double), ("Cv", C.c_double),
        ("Cp", C.c_double), ("mu", C.c_double), ("qk", C.c_double),
        ("S_over_T_ref", C.c_double),
        ("initvar", C.c_double * IBG_NFV),
        ("inlet_T0", C.c_double), ("inlet_p0", C.c_double),
        ("inlet_angle", C.c_double), ("outlet_p", C.c_double),
        ("cfl", C.c_double), ("dt_ratio", C.c_double), ("dt_init", C.c_double),
        ("force_ref", C.c_double), ("moment_ref", C.c_double),
        ("p_ref_area", C.c_double), ("shear_ref_area", C.c_double),
        ("nut_inf", C.c_double), ("nut_inlet", C.c_double), ("Pr_tur", C.c_double),
    ]


class BodyMotion(C.Structure):
    _fields_ = [("tra_vel", C.c_double * 3), ("rot_vel", C.c_double * 3)]


class Surface(C.Structure):
    _fields_ = [
        ("npoint", C.c_int32), ("pt_off", f64p), ("pat_pt0", i32p), ("pat_body", i32p),
        ("pat_flip", i8p), ("pat_dz", f64p), ("body_p0", i32p), ("case_origin", f64p),
        ("cell_patch", i32p), ("pat_close", i32p), ("ib_capacity", C.c_int32),
    ]


class BodyPose(C.Structure):
    _fields_ = [("pivot", C.c_double * 2), ("cosa", C.c_double), ("sina", C.c_double)]


class PitchHeave(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("aoa0_deg", C.c_double),
                ("amp_deg", C.c_double), ("omega", C.c_double), ("heave", C.c_double)]


class Config(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("nstage", C.c_int32), ("alpha", C.c_double * 4),
        ("dt_var_step", C.c_int32), ("temporal_refine", C.c_int32),
        ("limiter", C.c_int32), ("viscous", C.c_int32), ("tile", C.c_int32),
        ("turbulence", C.c_int32), ("schedule", C.c_int32), ("reflux", C.c_int32),
    ]


class Case(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dh0", C.c_double * 3), ("origin", C.c_double * 3),
        ("nlevels", C.c_int32), ("twod_from", C.c_int32),
        ("band_lo", C.c_double * IBG_MAX_LEVELS),
        ("band_hi", C.c_double * IBG_MAX_LEVELS),
        ("nbody", C.c_int32), ("body_npts", C.c_int32 * IBM_MAX_BODIES),
        ("body_xy", f64p), ("body_center", C.c_double * 3),
        ("span", C.c_double), ("span_patches", C.c_int32),
        ("gama", C.c_double), ("Ma", C.c_double), ("Re", C.c_double),
        ("Pr", C.c_double), ("cfl", C.c_double), ("backpressure", C.c_double),
        ("dt_init", C.c_double), ("face_bc", C.c_int32 * 4), ("refine_only", C.c_int32),
    ]


_libs = {}


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise OSError(
            f"{path} is not built; run `python __graft_entry__.py build` "
            "(or immsered-boundary_amd/build.py) first")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def host_lib():
    lib = _load("libibhost.so")
    if not getattr(lib, "_typed", False):
        lib.ibm_build.restype = C.c_void_p
        lib.ibm_build.argtypes = [C.POINTER(Case), C.c_int32]
        lib.ibm_mesh.restype = C.POINTER(MeshSoA)
        lib.ibm_mesh.argtypes = [C.c_void_p]
        lib.ibm_physics.restype = C.c_int32
        lib.ibm_physics.argtypes = [C.POINTER(Case), C.POINTER(Physics)]
        lib.ibm_init_state.restype = C.c_int32
        lib.ibm_init_state.argtypes = [C.c_void_p, C.POINTER(Physics), f64p]
        lib.ibm_cell_centers.restype = C.c_int32
        lib.ibm_cell_centers.argtypes = [C.c_void_p, f64p]
        lib.ibm_last_error.restype = C.c_char_p
        lib.ibm_free.restype = None
        lib.ibm_free.argtypes = [C.c_void_p]
        lib.ibm_naca4.restype = None
        lib.ibm_naca4.argtypes = [C.c_double, C.c_int32, C.c_double, C.c_double,
                                  C.c_double, f64p]
        lib.ibm_naca4_offsets.restype = None
        lib.ibm_naca4_offsets.argtypes = [C.c_double, C.c_int32, f64p]
        lib.ibm_pose.restype = None
        lib.ibm_pose.argtypes = [C.c_double, C.c_double, C.c_double, f64p]
        lib.ibm_transfer.restype = C.c_int32
        lib.ibm_transfer.argtypes = [C.c_void_p, C.POINTER(Physics), f64p, f64p, C.c_void_p, f64p,
                                     f64p]
        lib.ibm_write_grid.restype = C.c_int32
        lib.ibm_write_grid.argtypes = [C.c_void_p, C.c_char_p]
        lib.ibm_write_solution.restype = C.c_int32
        lib.ibm_write_solution.argtypes = [C.c_void_p, C.POINTER(Physics), f64p, f64p, C.c_int32,
                                           f64p, f64p, C.c_char_p]
        lib.ibm_read_solution.restype = C.c_int32
        lib.ibm_read_solution.argtypes = [C.c_void_p, C.POINTER(Physics), C.c_char_p, f64p, f64p,
                                          C.POINTER(C.c_int32), f64p, f64p]
        lib.ibm_surface.restype = C.c_int32
        lib.ibm_surface.argtypes = [C.c_void_p, f64p, C.c_int32, C.POINTER(Surface)]
        lib.ibm_cell_patch.restype = C.c_int32
        lib.ibm_cell_patch.argtypes = [C.c_void_p, i32p]
        lib.ibm_patch_close.restype = C.c_int32
        lib.ibm_patch_close.argtypes = [C.c_void_p, i32p]
        lib._typed = True
    return lib


GPU_FUNCS = {
    "ibg_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "ibg_destroy": (C.c_int, [C.c_void_p]),
    "ibg_version": (C.c_char_p, []),
    "ibg_check_mesh": (C.c_int, [C.POINTER(MeshSoA), C.POINTER(C.c_char_p)]),
    "ibg_upload_mesh": (C.c_int, [C.c_void_p, C.POINTER(MeshSoA), C.POINTER(Physics)]),
    "ibg_set_state": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_state": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_set_nut": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_nut": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_bc_fill": (C.c_int, [C.c_void_p, C.c_int32]),
    "ibg_interface_exchange": (C.c_int, [C.c_void_p, C.c_int32]),
    "ibg_ib_fill": (C.c_int, [C.c_void_p, C.POINTER(BodyMotion)]),
    "ibg_compute_dt": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "ibg_old_cons": (C.c_int, [C.c_void_p, C.c_int32]),
    "ibg_viscous_flux": (C.c_int, [C.c_void_p, C.c_int32]),
    "ibg_stage": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "ibg_time_march": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "ibg_prepare_march": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "ibg_residual": (C.c_int, [C.c_void_p, f64p, f64p]),
    "ibg_residual_nut": (C.c_int, [C.c_void_p, f64p, f64p]),
    "ibg_forces": (C.c_int, [C.c_void_p, C.POINTER(BodyMotion), f64p]),
    "ibg_get_faceflux": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_vsflux": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_source": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_mutflux": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_nut_source": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "ibg_get_dt": (C.c_int, [C.c_void_p, f64p]),
    "ibg_set_dt": (C.c_int, [C.c_void_p, f64p]),
    "ibg_last_error": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "ibg_synchronize": (C.c_int, [C.c_void_p]),
    "ibg_upload_surface": (C.c_int, [C.c_void_p, C.POINTER(MeshSoA), C.POINTER(Surface)]),
    "ibg_move_bodies": (C.c_int, [C.c_void_p, C.POINTER(BodyPose)]),
    "ibg_set_motion": (C.c_int, [C.c_void_p, C.POINTER(PitchHeave), C.c_double]),
    "ibg_get_time": (C.c_int, [C.c_void_p, f64p]),
    "ibg_set_time": (C.c_int, [C.c_void_p, f64p]),
    "ibg_get_geometry": (C.c_int, [C.c_void_p, f64p, i32p, i8p, i32p]),
    "ibg_get_ib_list": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, f64p, f64p, f64p, f64p,
                                  i32p]),
    "ibg_get_patches": (C.c_int, [C.c_void_p, f64p, f64p, f64p, i32p, f64p, f64p]),
    "ibg_last_timing": (C.c_int, [C.c_void_p, f64p, f64p, C.POINTER(C.c_int64)]),
    "ibg_bench_stage": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, f64p,
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ibg_bench_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, f64p]),
    "ibg_cell_stage_updates": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "ibg_autotune_tile": (C.c_int, [C.c_void_p, C.POINTER(MeshSoA), C.POINTER(Physics), i32p,
                                    C.c_int32, C.c_int32, i32p, f64p]),
    "ibg_closure_init": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64]),
    "ibg_closure_sample": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.POINTER(C.c_int64)]),
    "ibg_closure_fit_stats": (C.c_int, [C.c_void_p, C.c_double]),
    "ibg_closure_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_uint64, f64p]),
    "ibg_closure_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "ibg_closure_params": (C.c_int, [C.c_void_p, f32p, C.c_int64]),
    "ibg_closure_set_params": (C.c_int, [C.c_void_p, f32p, C.c_int64]),
    "ibg_closure_infer": (C.c_int, [C.c_void_p, f32p, f32p, C.c_int64]),
    "ibg_closure_infer_noalloc": (C.c_int, [C.c_void_p, f32p, f32p, C.c_int64]),
    "ibg_closure_grad": (C.c_int, [C.c_void_p, f32p, f32p, C.c_int64, f32p, f64p]),
    "ibg_closure_dataset": (C.c_int, [C.c_void_p, f32p, f32p, i64p, C.c_int64]),
    "ibg_closure_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
}


def gpu_lib():
    lib = _load("libibgpu.so")
    if not getattr(lib, "_typed", False):
        for name, (res, args) in GPU_FUNCS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._typed = True
    return lib
