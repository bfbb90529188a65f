"""ctypes binding of libmoons_amd.so (the C ABI declared in include/moons_amd.h).

The library is built in-tree by moons_amd.build; loading fails loudly when it
is missing -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmoons_amd.so")

c_float_p = C.POINTER(C.c_float)
c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)


class PcgInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("exit_loop", C.c_int),
        ("res_norm0", C.c_double),
        ("res_norm", C.c_double),
        ("ms", C.c_double),
    ]


# name -> (restype, argtypes); every symbol declared in include/moons_amd.h
class MhdParams(C.Structure):
    _fields_ = [("Re", C.c_double), ("Ha", C.c_double), ("Rem", C.c_double), ("dt", C.c_double),
                ("alpha_U", C.c_double), ("induction", C.c_int), ("prec_p", C.c_int), ("B0", C.c_double * 3),
                ("iter_max", C.c_int * 4), ("tol_rel", C.c_double * 4), ("tol_abs", C.c_double * 4),
                ("n_skip", C.c_int * 4), ("mpg_dir", C.c_int), ("mpg_scale", C.c_double),
                ("finite_Rem", C.c_int), ("prec_phi", C.c_int)]


class MhdInfo(C.Structure):
    _fields_ = [("iterations", C.c_int * 4), ("ms", C.c_double)]


class NrgParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("method", C.c_int), ("coeff_implicit", C.c_double), ("advection", C.c_int),
                ("advection_scale", C.c_double), ("diffusion", C.c_int), ("diffusion_scale", C.c_double),
                ("iter_max", C.c_int), ("tol_rel", C.c_double), ("tol_abs", C.c_double), ("n_skip", C.c_int),
                ("KE_diffusion", C.c_int), ("KE_diffusion_scale", C.c_double),
                ("viscous_dissipation", C.c_int), ("viscous_dissipation_scale", C.c_double),
                ("joule_heating", C.c_int), ("joule_heating_scale", C.c_double),
                ("volumetric_heating", C.c_int), ("volumetric_heating_scale", C.c_double),
                ("rk_stages", C.c_int)]


class NrgInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("ms", C.c_double)]


SIGNATURES = {
    "moons_init": (C.c_int, [C.c_int]),
    "moons_finalize": (C.c_int, []),
    "moons_last_error": (C.c_char_p, []),
    "moons_ktimer_start": (C.c_int, [C.c_int, C.c_int]),
    "moons_ktimer_stop": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "moons_synchronize": (C.c_int, []),
    "moons_stream": (C.c_void_p, []),
    "moons_nodes_uniform": (C.c_int, [C.c_double, C.c_double, C.c_int, c_double_p]),
    "moons_nodes_roberts_both": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_double, c_double_p]),
    "moons_mesh_create": (C.c_int, [c_double_p, C.c_int, c_double_p, C.c_int, c_double_p, C.c_int, C.POINTER(C.c_void_p)]),
    "moons_mesh_destroy": (C.c_int, [C.c_void_p]),
    "moons_mesh_coords": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, c_int_p]),
    "moons_field_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "moons_field_destroy": (C.c_int, [C.c_void_p]),
    "moons_field_shape": (C.c_int, [C.c_void_p, c_int_p]),
    "moons_field_set_bcs": (C.c_int, [C.c_void_p, c_int_p, c_double_p, C.c_int]),
    "moons_field_set_bc_props": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "moons_field_copy_bcs": (C.c_int, [C.c_void_p, C.c_void_p]),
    "moons_field_upload": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t]),
    "moons_field_download": (C.c_int, [C.c_void_p, c_double_p, C.c_size_t]),
    "moons_field_fill": (C.c_int, [C.c_void_p, C.c_double]),
    "moons_field_copy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "moons_field_device_ptr": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "moons_gf_write": (C.c_int, [C.c_char_p, c_double_p, c_int_p]),
    "moons_gf_read": (C.c_int, [C.c_char_p, c_double_p, C.c_size_t, c_int_p]),
    "moons_raw_write": (C.c_int, [c_double_p, C.c_int, c_double_p, C.c_int, c_double_p, C.c_int, C.c_int, c_double_p,
                                  C.c_char_p, C.c_char_p, C.c_int, c_double_p]),
    "moons_raw_read": (C.c_int, [c_int_p, C.c_int, c_double_p, C.c_char_p, C.c_char_p, C.c_int]),
    "moons_field_export": (C.c_int, [C.c_void_p, C.c_char_p]),
    "moons_field_import": (C.c_int, [C.c_void_p, C.c_char_p]),
    "moons_export_raw": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, c_double_p]),
    "moons_import_raw": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]),
    "moons_apply_bcs": (C.c_int, [C.c_void_p]),
    "moons_apply_bcs_implicit": (C.c_int, [C.c_void_p]),
    "moons_del": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "moons_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "moons_div": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "moons_curl": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "moons_lap_uniform_sf": (C.c_int, [C.c_void_p, C.c_void_p]),
    "moons_volume": (C.c_int, [C.c_void_p]),
    "moons_interp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "moons_edge_cross_face_E": (C.c_int, [C.POINTER(C.c_void_p)] * 4),
    "moons_advect_B": (C.c_int, [C.POINTER(C.c_void_p)] * 5),
    "moons_pcg_sf_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "moons_pcg_destroy": (C.c_int, [C.c_void_p]),
    "moons_pcg_set_isp": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]),
    "moons_pcg_sf_set_k": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "moons_pcg_sf_update_mfp": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "moons_nrg_create": (C.c_int, [C.c_void_p, C.POINTER(NrgParams), C.c_void_p, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "moons_nrg_destroy": (C.c_int, [C.c_void_p]),
    "moons_nrg_step": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(NrgInfo)]),
    "moons_nrg_set_joule": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p]),
    "moons_nrg_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "moons_pcg_sf_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PcgInfo)]),
    "moons_pcg_sf_solve_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int,
                                             C.POINTER(PcgInfo)]),
    "moons_clean_div": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p, C.c_double,
                                  C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(PcgInfo)]),
    "moons_pcg_time_kernels": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "moons_pcg_time_mg": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "moons_fft_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "moons_fft_destroy": (C.c_int, [C.c_void_p]),
    "moons_fft_set_coefficients": (C.c_int, [C.c_void_p, C.c_int]),
    "moons_fft_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "moons_clean_div_fft": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                      C.c_void_p]),
    "moons_pcg_set_mg": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "moons_pcg_mg_info": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "moons_pcg_apply_prec": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "moons_pcg_vf_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p)]),
    "moons_pcg_vf_destroy": (C.c_int, [C.c_void_p]),
    "moons_pcg_vf_set_isp": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]),
    "moons_pcg_vf_update_mfp": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "moons_pcg_vf_solve": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(PcgInfo)]),
    "moons_mhd_create": (C.c_int, [C.c_void_p, C.POINTER(MhdParams), C.POINTER(C.c_void_p), C.c_void_p,
                                   C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(C.c_void_p)]),
    "moons_mhd_create_md": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(MhdParams), C.POINTER(C.c_void_p),
                                      C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p)]),
    "moons_nodes_ext_uniform": (C.c_int, [c_double_p, C.c_int, C.c_int, c_double_p]),
    "moons_nodes_ext": (C.c_int, [c_double_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, c_double_p,
                                  C.c_int, C.POINTER(C.c_int)]),
    "moons_domain_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "moons_domain_destroy": (C.c_int, [C.c_void_p]),
    "moons_domain_overlap": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_int_p]),
    "moons_embed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "moons_extract": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "moons_set_sigma_inv": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
    "moons_mhd_destroy": (C.c_int, [C.c_void_p]),
    "moons_mhd_step": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(MhdInfo)]),
    "moons_mhd_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "moons_mhd_set_buoyancy": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_double,
                                         C.c_int, C.c_double]),
    "moons_mhd_graph_capture": (C.c_int, [C.c_void_p]),
    "moons_mhd_graph_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "moons_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "moons_stream_destroy": (C.c_int, [C.c_void_p]),
    "moons_stream_synchronize": (C.c_int, [C.c_void_p]),
    "moons_unet_create": (C.c_int, [C.c_ulonglong, C.POINTER(C.c_void_p)]),
    "moons_unet_destroy": (C.c_int, [C.c_void_p]),
    "moons_unet_num_params": (C.c_size_t, [C.c_void_p]),
    "moons_unet_get_weights": (C.c_int, [C.c_void_p, c_float_p, C.c_size_t]),
    "moons_unet_set_weights": (C.c_int, [C.c_void_p, c_float_p, C.c_size_t]),
    "moons_unet_forward_host": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_float_p, c_float_p]),
    "moons_unet_grad_host": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_float_p, c_float_p,
                                       c_double_p, c_float_p]),
    "moons_unet_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, c_float_p, c_float_p, C.c_int,
                                   C.c_int, C.c_double, C.c_ulonglong, c_double_p]),
    "moons_mfma_check": (C.c_int, [c_float_p, c_float_p, c_float_p]),
    "moons_mfma_check32": (C.c_int, [c_float_p, c_float_p, c_float_p]),
    "moons_pcg_set_unet": (C.c_int, [C.c_void_p, C.c_void_p]),
    "moons_pcg_set_mg_unet": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "moons_pcg_set_mg_unet_cycles": (C.c_int, [C.c_void_p, C.c_int]),
    "moons_pcg_set_mg_gamma": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "moons_pcg_set_fallback": (C.c_int, [C.c_void_p, C.c_int]),
}

_lib = None


class MoonsError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libmoons_amd.so (built in-tree); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MoonsError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP library has no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().moons_last_error().decode(errors="replace")
        raise MoonsError(f"libmoons_amd status {status}: {msg}")
