"""ctypes binding of ``libjaxiga_hip.so`` (the C ABI in ``include/jaxiga_hip.h``).

The library is required: every compute entry point of the package goes
through it, and a missing or unloadable library raises immediately -- there is
no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libjaxiga_hip.so"

c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P_int = ctypes.POINTER(ctypes.c_int32)
P_dbl = ctypes.POINTER(ctypes.c_double)

IGA_POISSON, IGA_ELASTICITY = 0, 1
IGA_COEF_CONST, IGA_COEF_POINTS = 0, 1
IGA_SRC_NONE, IGA_SRC_CONST, IGA_SRC_POINTS = 0, 1, 2
IGA_FP64, IGA_FP32, IGA_BF16 = 0, 1, 2
IGA_OK, IGA_ERR_SHAPE, IGA_ERR_UNSUPPORTED, IGA_ERR_HIP, IGA_ERR_ARG = 0, -1, -2, -3, -4
IGA_MF_RATIONAL, IGA_MF_SEPARABLE, IGA_MF_ELASTICITY = 1, 2, 4

STATUS = {-1: "inconsistent shapes", -2: "unsupported degree/dimension/quadrature",
          -3: "HIP runtime error", -4: "null pointer or invalid argument"}


class IgaPatch(ctypes.Structure):
    _fields_ = [
        ("dim", c_int), ("degree", c_int), ("n_elem", c_int * 3), ("n_cp", c_int * 3),
        ("ext", c_void_p * 3), ("first", c_void_p * 3), ("lo", c_void_p * 3), ("hi", c_void_p * 3),
        ("cpts", c_void_p), ("wgts", c_void_p), ("elo", c_void_p * 3), ("ehi", c_void_p * 3),
        ("regular", c_int),
    ]


class IgaRule(ctypes.Structure):
    _fields_ = [("n_q1", c_int), ("pts", c_void_p), ("wts", c_void_p)]


class IgaAsmArgs(ctypes.Structure):
    _fields_ = [
        ("problem", c_int), ("vec", c_int), ("precision", c_int), ("coef_kind", c_int),
        ("coef_const", c_double), ("coef_points", c_void_p), ("cmat", c_double * 36),
        ("src_kind", c_int), ("src_const", c_double * 3), ("src_points", c_void_p),
        ("full_to_free", c_void_p), ("freed", c_void_p), ("gfull", c_void_p), ("nnz", c_i64),
        ("rows", c_void_p), ("cols", c_void_p), ("vals", c_void_p), ("ke_work", c_void_p),
        ("ke_work_bytes", c_i64), ("rhs", c_void_p), ("diag", c_void_p), ("passes", c_int),
        ("batch", c_int), ("n_free", c_int), ("indptr", c_void_p), ("ke_full", c_int),
    ]


class IgaNet(ctypes.Structure):
    _fields_ = [
        ("n_layers", c_int), ("kind", c_int * 16), ("cin", c_int * 16), ("cout", c_int * 16),
        ("stride", c_int * 16), ("src", c_int * 16), ("skip", c_int * 16), ("woff", c_i64 * 16),
        ("H", c_int * 17), ("W", c_int * 17), ("c_in", c_int), ("relu", c_int * 16),
        ("boff", c_i64 * 16), ("precision", c_int),
    ]


class IgaMgLevel(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("h", c_int), ("w", c_int), ("nnz", c_int), ("indptr", c_void_p),
        ("cols", c_void_p), ("vals", c_void_p), ("dsq", c_void_p), ("stencil", c_void_p),
        ("omega", ctypes.c_float), ("p_nnz", c_int), ("pt_nnz", c_int), ("p_ptr", c_void_p),
        ("p_cols", c_void_p), ("p_vals", c_void_p), ("pt_ptr", c_void_p), ("pt_cols", c_void_p),
        ("pt_vals", c_void_p), ("stencil16", c_void_p),
    ]


class IgaMgSetupLevel(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("h", c_int), ("w", c_int), ("nnz", c_int), ("indptr", c_void_p),
        ("cols", c_void_p), ("vals64", c_void_p), ("vals32", c_void_p), ("dinv", c_void_p),
        ("dsq", c_void_p), ("pt_ptr", c_void_p), ("pt_cols", c_void_p), ("pt_vals", c_void_p),
    ]


class IgaMgSetupGrid(ctypes.Structure):
    _fields_ = [("f_lo", (c_void_p * 2) * 16), ("rw", (c_void_p * 2) * 16), ("op64", c_void_p * 16),
                ("op32", c_void_p * 16), ("feats", c_void_p), ("work", c_void_p),
                ("tail_level", c_int), ("omega", c_double), ("tail", c_void_p), ("tail_work", c_void_p),
                ("regular", c_int * 16), ("stages", c_int), ("gal_multi", c_int), ("gal_src", c_int * 16),
                ("gal_f_lo", (c_void_p * 2) * 16), ("gal_rw", (c_void_p * 2) * 16), ("gal_work", c_void_p),
                ("vals0_slot", c_void_p)]


class IgaMgSetup(ctypes.Structure):
    _fields_ = [("n_levels", c_int), ("radius", c_int), ("level", IgaMgSetupLevel * 16),
                ("coarse_inv", c_void_p), ("inv_work", c_void_p),
                ("grid", ctypes.POINTER(IgaMgSetupGrid))]


class IgaMgBand(ctypes.Structure):
    _fields_ = [("c_lo", c_void_p), ("pw", c_void_p), ("f_lo", c_void_p), ("rw", c_void_p)]


class IgaMgGrid(ctypes.Structure):
    _fields_ = [("n_levels", c_int), ("radius", c_int), ("op", c_void_p * 16),
                ("band", (IgaMgBand * 2) * 16), ("op64", c_void_p), ("feats", c_void_p)]


class IgaMg(ctypes.Structure):
    _fields_ = [("n_levels", c_int), ("ks", c_int), ("level", IgaMgLevel * 16),
                ("coarse_inv", c_void_p), ("work", c_void_p), ("tail_level", c_int),
                ("tail", c_void_p), ("grid", ctypes.POINTER(IgaMgGrid))]


_SIGNATURES = {
    "iga_assemble": (c_int, [ctypes.POINTER(IgaPatch), ctypes.POINTER(IgaRule),
                             ctypes.POINTER(IgaAsmArgs), c_void_p]),
    "iga_assemble_workspace": (c_i64, [ctypes.POINTER(IgaPatch), c_int, c_int]),
    "iga_basis_eval": (c_int, [ctypes.POINTER(IgaPatch), c_int, c_void_p, c_int, c_void_p, c_int,
                               c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p]),
    "iga_load_points": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_csr_count": (c_int, [c_int, P_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                              c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_csr_fill": (c_int, [c_int, P_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                             c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p]),
    "iga_scan_workspace": (c_i64, [c_int]),
    "iga_scan_exclusive": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_csr_diag_inv": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_spmv": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_cg_workspace": (c_i64, [c_int]),
    "iga_cg_solve": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                             c_void_p, c_double, c_int, c_void_p, P_int, P_dbl, P_dbl, c_void_p]),
    "iga_cg1_solve": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                              c_void_p, c_double, c_int, c_void_p, P_int, P_dbl, P_dbl, c_void_p]),
    "iga_cg_solve_noalloc": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                     c_void_p, c_double, c_int, P_int, P_dbl, P_dbl, c_void_p]),
    "iga_cg_solve_mixed": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_double, c_int, c_int, c_void_p, P_int,
                                   P_dbl, P_dbl, c_void_p]),
    "iga_residual_vec": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "iga_axpy": (c_int, [c_int, c_double, c_void_p, c_void_p, c_void_p]),
    "iga_residual_norms": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "iga_mf_workspace": (c_i64, [ctypes.POINTER(IgaPatch), c_int, c_int]),
    "iga_mf_setup": (c_int, [ctypes.POINTER(IgaPatch), ctypes.POINTER(IgaRule), c_int, c_int, c_int,
                             c_double, c_void_p, c_void_p, c_i64, c_void_p]),
    "iga_mf_setup_elasticity": (c_int, [ctypes.POINTER(IgaPatch), ctypes.POINTER(IgaRule), c_int, c_int,
                                        c_void_p, c_void_p, c_i64, c_void_p]),
    "iga_mf_apply": (c_int, [ctypes.POINTER(IgaPatch), c_void_p, c_i64, c_int, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p]),
    "iga_mf_diag": (c_int, [ctypes.POINTER(IgaPatch), c_void_p, c_i64, c_int, c_void_p, c_void_p,
                            c_void_p, c_void_p]),
    "iga_cg_solve_mf": (c_int, [ctypes.POINTER(IgaPatch), c_void_p, c_i64, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_double, c_int, c_void_p, P_int, P_dbl,
                                P_dbl, c_void_p]),
    "iga_diff_norms": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_inv_guard": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "iga_csr_outer": (c_int, [ctypes.c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                              c_void_p]),
    "iga_elem_pattern_workspace": (c_i64, [c_int, c_int, c_int]),
    "iga_elem_pattern": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_i64,
                                 c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_i64), c_void_p]),
    "iga_elem_basis": (c_int, [c_int, c_int, c_int, c_i64, c_i64, ctypes.POINTER(IgaRule), c_void_p,
                               c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p]),
    "iga_elem_assemble": (c_int, [c_int, c_int, c_int, c_i64, c_i64, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_int, ctypes.POINTER(IgaAsmArgs), c_void_p, c_void_p]),
    "iga_neohooke": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p, c_void_p,
                             c_void_p]),
    "iga_point_forces": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "iga_field_points": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_net_act_floats": (c_i64, [ctypes.POINTER(IgaNet), c_int]),
    "iga_net_act_offset": (c_i64, [ctypes.POINTER(IgaNet), c_int, c_int]),
    "iga_net_scratch_floats": (c_i64, [ctypes.POINTER(IgaNet), c_int]),
    "iga_net_forward": (c_int, [ctypes.POINTER(IgaNet), c_int, c_void_p, c_void_p, c_void_p]),
    "iga_net_infer": (c_int, [ctypes.POINTER(IgaNet), c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p]),
    "iga_net_pack16": (c_int, [ctypes.POINTER(IgaNet), c_int, c_void_p, c_void_p, c_void_p]),
    "iga_net_backward": (c_int, [ctypes.POINTER(IgaNet), c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p]),
    "iga_nmse": (c_int, [c_int, c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_adam": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, ctypes.c_float, c_int, c_void_p]),
    "iga_mg_load": (c_int, [ctypes.POINTER(IgaMgSetup), c_void_p, c_void_p]),
    "iga_mg_work_floats": (c_i64, [ctypes.POINTER(IgaMg)]),
    "iga_mg_vcycle": (c_int, [ctypes.POINTER(IgaMg), c_int, c_void_p, c_void_p, c_void_p]),
    "iga_mg_dense_ld": (c_int, [c_int]),
    "iga_mg_tail_dense_doubles": (c_i64, [ctypes.POINTER(IgaMgSetup), c_int]),
    "iga_mg_grid_ops": (c_int, [ctypes.POINTER(IgaMg), c_void_p, c_void_p]),
    "iga_mg_tail_floats": (c_i64, [ctypes.POINTER(IgaMg), c_int]),
    "iga_mg_tail": (c_int, [ctypes.POINTER(IgaMg), c_int, c_void_p, c_void_p, c_void_p]),
    "iga_masked_spgemm": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_stencil_features": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "iga_dense_inverse": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "iga_csr_f32": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_int, c_int, c_i64, c_i64, c_void_p]),
    "iga_stencil_apply": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                  c_int, c_int, c_void_p]),
    "iga_stencil_wgrad_t": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    "iga_energy_loss": (c_int, [c_int, c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p]),
    "iga_fcg_workspace": (c_i64, [c_int]),
    "iga_fcg_mg_plan": (c_int, [ctypes.POINTER(IgaMg), c_int, c_int, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                ctypes.POINTER(c_void_p)]),
    "iga_fcg_mg_run": (c_int, [c_void_p, c_double, c_int, P_int, P_dbl, P_dbl, P_dbl, c_void_p,
                               c_void_p, c_void_p]),
    "iga_fcg_mg_destroy": (None, [c_void_p]),
    "iga_fcg_mg": (c_int, [ctypes.POINTER(IgaMg), c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_double, c_int, c_int, c_void_p, c_void_p, P_int, P_dbl, P_dbl,
                           c_void_p]),
    "iga_batched_cg": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_double, c_int, c_void_p,
                               c_void_p, c_void_p]),
    "iga_batched_cg_seg_plan": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "iga_batched_cg_seg_workspace": (c_i64, [c_int, c_int]),
    "iga_batched_cg_seg": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                   c_double, c_int, c_void_p, c_i64, c_void_p, c_void_p, c_void_p]),
    "iga_rff_eval": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_double,
                             c_void_p, c_void_p]),
}

_LIB = None


def exported_symbols():
    return sorted(_SIGNATURES)


def load():
    """Load and type the library; raises ``RuntimeError`` if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = Path(os.environ.get("JAXIGA_HIP_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(
            f"jaxiga_amd: HIP library {path} is missing; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)"
        )
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {STATUS.get(rc, 'status')} ({rc})")


def ptr(t):
    """Device (or host) address of a tensor, or None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
