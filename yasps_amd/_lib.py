"""ctypes binding of ``libyasps_amd.so`` (the C ABI declared in ``include/yasps_amd.h``).

The library is required: importing the kernels without it raises, there is no CPU
fallback for any device entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libyasps_amd.so"

MAX_UNION_CHILDREN = 8
MAX_PROGRAM_CONNS = 8
MAX_PROGRAM_UNIONS = 8


class UnionF64(C.Structure):
    _fields_ = [("num_children", C.c_uint32), ("width", C.c_uint32),
                ("counts", C.c_void_p), ("data", C.c_void_p * MAX_UNION_CHILDREN)]


class ScatterArgs(C.Structure):
    _fields_ = [("segment_indices", C.c_void_p), ("segment_sizes", C.c_void_p),
                ("local_permutations", C.c_void_p), ("lookups", C.c_void_p),
                ("coordinates_outer", C.c_void_p), ("max_num_indices", C.c_uint32),
                ("num_attributes", C.c_uint32), ("diagonal_blocks_start", C.c_void_p),
                ("gradient_segments_start", C.c_void_p), ("gradient", C.c_void_p),
                ("hessian_blocks", C.c_void_p), ("diagonal_blocks", C.c_void_p),
                ("block_staging", C.c_void_p), ("gradient_staging", C.c_void_p),
                ("gradient_local_size", C.c_uint32)]


class AffineChildren(C.Structure):
    _fields_ = [("mask", C.c_uint32), ("vertex_body", C.c_void_p * MAX_UNION_CHILDREN),
                ("rest_position", C.c_void_p * MAX_UNION_CHILDREN)]


class Pcg3System(C.Structure):
    _fields_ = [("row_ptr", C.c_void_p), ("col", C.c_void_p), ("values", C.c_void_p),
                ("nnz", C.c_uint32), ("diagonal_inverse", C.c_void_p), ("rhs", C.c_void_p),
                ("n", C.c_uint32), ("solution", C.c_void_p), ("work", C.c_void_p),
                ("cg_scalars", C.c_void_p), ("max_iterations", C.c_uint32),
                ("passes_done", C.c_uint32), ("status", C.c_int32), ("iterations", C.c_int32),
                ("delta0", C.c_double), ("delta", C.c_double), ("rr", C.c_double),
                ("mas", C.c_void_p)]


PCG3_RUNNING = 0x7FFFFFFF


class BlockMatrix(C.Structure):
    _fields_ = [("values", C.c_void_p), ("positions", C.c_void_p),
                ("num_categories", C.c_uint32), ("starts", C.POINTER(C.c_uint32)),
                ("counts", C.POINTER(C.c_uint32)), ("dims", C.POINTER(C.c_uint32))]


class CoordDim(C.Structure):
    _fields_ = [("row", C.c_uint32), ("col", C.c_uint32), ("h", C.c_uint16),
                ("w", C.c_uint16), ("placeholder", C.c_uint32)]


VALUES_F64, VALUES_F32, VALUES_UPPER = 0, 1, 2


class UpperValues(C.Structure):
    """ysp_upper_values: the upper block storage read through each entry's source word."""
    _fields_ = [("src", C.c_void_p), ("src_64", C.c_int32), ("static_values", C.c_void_p),
                ("dynamic_values", C.c_void_p)]


class PcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("status", C.c_int32), ("delta0", C.c_double),
                ("delta", C.c_double), ("rr0", C.c_double), ("rr", C.c_double)]


MAS_MAX_COARSE_LEVELS = 8


class MasArgs(C.Structure):
    _fields_ = [("restrict_offsets", C.c_void_p), ("restrict_items", C.c_void_p),
                ("num_coarse_nodes", C.c_uint32), ("num_coarse_levels", C.c_uint32),
                ("level_offsets", C.c_uint32 * (MAS_MAX_COARSE_LEVELS + 1)),
                ("prolong_maps", C.c_void_p), ("inverses", C.c_void_p),
                ("domain_nodes", C.c_void_p), ("domain_node_offsets", C.c_void_p),
                ("num_domains", C.c_uint32), ("level0_domains", C.c_uint32),
                ("weights", C.c_void_p), ("coarse_residual", C.c_void_p),
                ("fine_correction", C.c_void_p), ("coarse_correction", C.c_void_p),
                ("inverses_bf16", C.c_void_p), ("bf16_first_domain", C.c_uint32),
                ("domain_scale", C.c_void_p), ("domain_size", C.c_uint32)]


P = C.c_void_p
U32 = C.c_uint32
I32 = C.c_int32
D = C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "ysp_last_error": (C.c_char_p, []),
    "ysp_version": (C.c_int, []),
    "ysp_sync": (C.c_int, [P]),
    "ysp_check_deferred": (C.c_int, [P]),
    "ysp_release_scratch": (None, []),
    "ysp_get_indices": (C.c_int, [P, I32, P, I32, P, I32, P, P, P, U32, U32, P]),
    "ysp_compress_indices": (C.c_int, [P, P, P, P, P, P, P, P, P, U32, U32, U32, P]),
    "ysp_compute_coordinates": (C.c_int, [P, P, P, P, P, P, U32, U32, P]),
    "ysp_get_unique_coords": (C.c_int, [P, P, U32, P, P, C.POINTER(U32), P]),
    "ysp_compress_unique_coords": (C.c_int, [P, U32, U32, P, P, P, P, P, C.POINTER(U32), P]),
    "ysp_snh_rest_data": (C.c_int, [C.POINTER(UnionF64), P, U32, P, P, P]),
    "ysp_eval_snh_tet": (C.c_int, [C.POINTER(UnionF64), P, U32, P, P, P, U32, P, U32, D, U32,
                                   C.POINTER(ScatterArgs), P]),
    "ysp_eval_inertia": (C.c_int, [P, P, P, P, D, U32, U32, C.POINTER(ScatterArgs), P]),
    "ysp_energy_snh_tet": (C.c_int, [C.POINTER(UnionF64), P, U32, P, P, P, U32, P, U32, D, P, P]),
    "ysp_energy_inertia": (C.c_int, [P, P, P, P, D, U32, P, P]),
    "ysp_snh_theta_adjoint": (C.c_int, [C.POINTER(UnionF64), P, U32, P, P, P, U32, P, U32, D,
                                        C.POINTER(ScatterArgs), P, P, P]),
    "ysp_energy_arity": (C.c_int, [I32, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32),
                                   C.POINTER(U32)]),
    "ysp_element_rest_data": (C.c_int, [I32, C.POINTER(UnionF64), P, U32, D, P, P]),
    "ysp_eval_element_energy": (C.c_int, [I32, C.POINTER(UnionF64), C.POINTER(AffineChildren), P,
                                          U32, P, P, U32, C.POINTER(ScatterArgs), P]),
    "ysp_affine_positions": (C.c_int, [P, P, P, P, U32, P, P]),
    "ysp_element_energy_values": (C.c_int, [I32, C.POINTER(UnionF64), P, U32, P, P, P, P]),
    "ysp_element_param_jacobian": (C.c_int, [I32, C.POINTER(UnionF64), P, U32, P, P, U32, P, U32,
                                             P, U32, U32, P, P]),
    "ysp_jacobian_rows": (C.c_int, [P, P, U32, U32, U32, P, P]),
    "ysp_jacobian_spmv": (C.c_int, [P, U32, U32, U32, P, I32, P, P, I32, P]),
    "ysp_block_spmv": (C.c_int, [C.POINTER(BlockMatrix), C.POINTER(BlockMatrix), P, P, U32,
                                 C.c_int, C.c_int, P]),
    "ysp_invert_diagonal_blocks": (C.c_int, [P, P, C.POINTER(U32), C.POINTER(U32),
                                             C.POINTER(U32), I32, P]),
    "ysp_diagonal_from_blocks": (C.c_int, [P, P, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32),
                                           C.POINTER(U32), I32, P]),
    "ysp_pcg_solve": (C.c_int, [U32, D, I32, C.POINTER(BlockMatrix), C.POINTER(BlockMatrix), P,
                                C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), C.POINTER(U32),
                                I32, P, U32, P, P, I32, P, P, C.POINTER(PcgStats), P]),
    "ysp_pcg_solve_full": (C.c_int, [U32, D, I32, P, P, P, I32, U32, U32, P, C.POINTER(U32),
                                     C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), I32, P, U32,
                                     P, P, I32, P, P, C.POINTER(PcgStats), P]),
    "ysp_bsr3_build_split": (C.c_int, [C.POINTER(BlockMatrix), C.POINTER(BlockMatrix), U32, P, P,
                                       P, U32, C.POINTER(U32), P]),
    "ysp_bsr3_gather_split": (C.c_int, [P, P, P, U32, P, P]),
    "ysp_inverse_map": (C.c_int, [P, P, U32, U32, P, P, P]),
    "ysp_gradient_contributions": (C.c_int, [P, P, U32, U32, C.c_uint64, U32, P, P, P]),
    "ysp_gather_blocks3": (C.c_int, [P, P, P, U32, P, P, P, P, P, U32, I32, P]),
    "ysp_gather_gradient3": (C.c_int, [P, P, P, U32, P, P]),
    "ysp_bsr3_build": (C.c_int, [P, U32, P, U32, U32, P, P, P, U32, C.POINTER(U32), P]),
    "ysp_bsr3_gather_values": (C.c_int, [P, P, P, U32, P, P, P]),
    "ysp_cast_f32": (C.c_int, [P, P, C.c_uint64, P]),
    "ysp_bsr3_spmv": (C.c_int, [P, P, P, I32, U32, U32, P, P, P]),
    "ysp_bsr3_refine": (C.c_int, [P, P, P, I32, U32, U32, P, P, P, P, C.POINTER(D), C.POINTER(D),
                                  P]),
    "ysp_pcg3_solve": (C.c_int, [U32, D, I32, P, P, P, I32, U32, P, I32, P, U32, P, P, I32, P, P,
                                 C.POINTER(PcgStats), P]),
    "ysp_pcg3_product_pass": (C.c_int, [P, P, P, I32, U32, U32, P, P, I32, P]),
    "ysp_pcg3_solve_noalloc": (C.c_int, [U32, D, I32, P, P, P, I32, U32, P, I32, P, U32, P, P,
                                         I32, C.POINTER(PcgStats), P]),
    "ysp_pcg3_batch_start": (C.c_int, [C.POINTER(Pcg3System), D, I32, I32, P]),
    "ysp_pcg3_batch_run": (C.c_int, [C.POINTER(Pcg3System), U32, U32, I32, I32, P]),
    "ysp_ccd_candidates": (C.c_int, [P, P, D, U32, P, U32, P, U32, P, U32, P, D, P, U32,
                                     C.POINTER(U32), P]),
    "ysp_ccd_active_pairs": (C.c_int, [P, P, P, P, U32, D, P, P, P, P, U32, C.POINTER(U32), P]),
    "ysp_ccd_step_bound": (C.c_int, [P, P, P, P, P, U32, D, P, C.POINTER(D), P]),
    "ysp_mas_coarsen": (C.c_int, [P, I32, U32, P, P, U32, P, P]),
    "ysp_mas_domain_scatter": (C.c_int, [P, I32, U32, P, P, U32, P]),
    "ysp_mas_invert_domains": (C.c_int, [P, P, U32, P, P, U32, P, U32, P]),
    "ysp_mas_apply": (C.c_int, [C.POINTER(MasArgs), P, U32, P, P]),
    "ysp_pcg3_mas_solve": (C.c_int, [U32, D, I32, P, P, P, I32, U32, C.POINTER(MasArgs), P, U32,
                                     P, P, I32, P, P, C.POINTER(PcgStats), P]),
}

_lib = None


def load(path: os.PathLike | None = None):
    """Load the shared library once and attach the signatures. Raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    # PyTorch-ROCm ships its own HIP runtime whose soname (libamdhip64.so.7) is the one
    # the library needs: importing torch first makes the dynamic loader reuse that
    # runtime instead of loading a second one from /opt/rocm.
    import torch  # noqa: F401
    if not p.exists():
        raise ImportError(
            f"yasps_amd: {p} is missing. Build it with `python -m yasps_amd.build` "
            "(hipcc, gfx950); there is no CPU fallback.")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class YspError(RuntimeError):
    pass


def check(rc: int, what: str) -> int:
    if rc < 0 and rc in (-1, -2, -6):
        msg = load().ysp_last_error()
        raise YspError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def u32_array(values):
    arr = (C.c_uint32 * max(1, len(values)))(*[int(v) for v in values])
    return arr
