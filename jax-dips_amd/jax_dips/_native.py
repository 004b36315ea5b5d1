"""ctypes binding of libdips_hip.so (the C-ABI declared in include/dips_hip.h).

The library is loaded from the package directory.  There is no fallback: if the
library is missing or fails to load, every entry point raises.  Device buffers
are passed as raw pointers (torch tensors are used only to own device memory
and to name the current HIP stream).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdips_hip.so")

DIPS_OK = 0


class DipsError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    _fields_ = [("xs", ctypes.c_void_p), ("ys", ctypes.c_void_p), ("zs", ctypes.c_void_p),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("extra", ctypes.c_void_p), ("nextra", ctypes.c_int64)]


class Records(ctypes.Structure):
    _fields_ = [("row_point", ctypes.c_void_p), ("row_ptr", ctypes.c_void_p), ("row_c", ctypes.c_void_p),
                ("ent_code", ctypes.c_void_p), ("ent_w", ctypes.c_void_p), ("nrows", ctypes.c_int64),
                ("nnz", ctypes.c_int64), ("hx", ctypes.c_float), ("hy", ctypes.c_float), ("hz", ctypes.c_float),
                ("row_coeffs", ctypes.c_void_p)]


class Net(ctypes.Structure):
    _fields_ = [("Lm", ctypes.c_int32), ("Hm", ctypes.c_int32), ("Lp", ctypes.c_int32), ("Hp", ctypes.c_int32),
                ("pc_w1", ctypes.c_int32), ("pc_w2", ctypes.c_int32), ("pc_scale", ctypes.c_float),
                ("act", ctypes.c_int32)]


class Opt(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("sched", ctypes.c_int32), ("lr", ctypes.c_float),
                ("decay_rate", ctypes.c_float), ("transition_steps", ctypes.c_float), ("max_norm", ctypes.c_float)]


class MfmaBatch(ctypes.Structure):
    _fields_ = [("row_begin", ctypes.c_int64), ("nrows", ctypes.c_int64), ("ent_begin", ctypes.c_int64),
                ("ent_end", ctypes.c_int64), ("t_row", ctypes.c_void_p), ("t_code", ctypes.c_void_p),
                ("t_w", ctypes.c_void_p), ("t_csr", ctypes.c_void_p), ("t_begin", ctypes.c_int64),
                ("n_plus", ctypes.c_int64), ("n_minus", ctypes.c_int64), ("t_count", ctypes.c_int64),
                ("t_group", ctypes.c_void_p), ("g_begin", ctypes.c_int64), ("g_plus", ctypes.c_int64),
                ("g_minus", ctypes.c_int64), ("i_pos", ctypes.c_void_p), ("i_count", ctypes.c_int64),
                ("e_item", ctypes.c_void_p)]


class Tiles(ctypes.Structure):
    _fields_ = [("tile_row", ctypes.c_void_p), ("nbounds", ctypes.c_int64), ("slot_x", ctypes.c_void_p),
                ("slot_y", ctypes.c_void_p), ("slot_z", ctypes.c_void_p), ("slot_w", ctypes.c_void_p),
                ("slot_c", ctypes.c_void_p), ("slot_info", ctypes.c_void_p), ("slot_row", ctypes.c_void_p)]


class BatchProblem(ctypes.Structure):
    _fields_ = [("grid", Grid), ("rec", Records), ("tiles", Tiles), ("batch_tile_begin", ctypes.c_void_p),
                ("nbatches", ctypes.c_int32), ("batch_size", ctypes.c_int64), ("params", ctypes.c_void_p),
                ("m", ctypes.c_void_p), ("v", ctypes.c_void_p), ("count", ctypes.c_void_p),
                ("loss_hist", ctypes.c_void_p), ("loss_cap", ctypes.c_int32)]


class BatchCfg(ctypes.Structure):
    _fields_ = [("blk_cap", ctypes.c_int32), ("grid_blocks", ctypes.c_int32)]


class BatchPlan(ctypes.Structure):
    _fields_ = [("nprob", ctypes.c_int32), ("nitems", ctypes.c_int32), ("blk_cap", ctypes.c_int32),
                ("grid_blocks", ctypes.c_int32), ("max_steps", ctypes.c_int32), ("ctrl_units", ctypes.c_int32)]


class HashGrid(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("features", ctypes.c_int32), ("table_size", ctypes.c_int64),
                ("n_min", ctypes.c_int32), ("n_max", ctypes.c_int32), ("bound", ctypes.c_float),
                ("encoder", ctypes.c_int32)]


class Assemble(ctypes.Structure):
    _fields_ = [("xs", ctypes.c_void_p), ("ys", ctypes.c_void_p), ("zs", ctypes.c_void_p),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("hx", ctypes.c_float), ("hy", ctypes.c_float), ("hz", ctypes.c_float),
                ("bnd", ctypes.c_float * 6),
                ("phis", ctypes.c_void_p), ("region", ctypes.c_void_p), ("cc_slot", ctypes.c_void_p),
                ("vol_m", ctypes.c_void_p), ("area_m", ctypes.c_void_p), ("tri_a", ctypes.c_void_p),
                ("beta_tri", ctypes.c_void_p), ("reserved0", ctypes.c_int32),
                ("mu_face", ctypes.c_void_p), ("pt_vals", ctypes.c_void_p), ("cs_slot", ctypes.c_void_p),
                ("cs_a", ctypes.c_void_p), ("cs_k", ctypes.c_void_p), ("cs_sel", ctypes.c_void_p),
                ("cs_delta", ctypes.c_void_p), ("cs_vals", ctypes.c_void_p), ("phi_cube", ctypes.c_void_p),
                ("scratch", ctypes.c_void_p), ("ip_slot", ctypes.c_void_p), ("mode", ctypes.c_int32),
                ("nnz", ctypes.c_void_p), ("ent_off", ctypes.c_void_p), ("row_of", ctypes.c_void_p),
                ("ent_code", ctypes.c_void_p), ("ent_w", ctypes.c_void_p), ("row_c", ctypes.c_void_p),
                ("npts", ctypes.c_int64), ("extra", ctypes.c_void_p), ("nextra", ctypes.c_int64),
                ("row_coeffs", ctypes.c_void_p)]


# name -> (restype, argtypes); must match include/dips_hip.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
SIGNATURES = {
    "dips_last_error": (ctypes.c_char_p, []),
    "dips_version": (ctypes.c_int, []),
    "dips_param_count": (_I64, [ctypes.POINTER(Net)]),
    "dips_partial_size": (_I64, [ctypes.POINTER(Net), _I64]),
    "dips_net_supported": (ctypes.c_int, [ctypes.POINTER(Net)]),
    "dips_loss_grad": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), _I64, _I64, _I64,
                                      ctypes.POINTER(Net), _P, _P, _I64, ctypes.c_int, ctypes.POINTER(_I32), _P]),
    "dips_opt_step": (ctypes.c_int, [ctypes.POINTER(Net), ctypes.POINTER(Opt), _P, _I32, ctypes.c_float, _P, _P, _P,
                                     _P, _P, _I32, _P]),
    "dips_train_level": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(_I64),
                                        ctypes.POINTER(_I64), ctypes.POINTER(_I64), _I32, _I64, _I32,
                                        ctypes.POINTER(Net),
                                        ctypes.POINTER(Opt), _P, _P, _P, _P, _P, _I32, _P, _I64, _P, _P]),
    "dips_train_step": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), _I64, _I64, _I64,
                                       ctypes.POINTER(Net),
                                       ctypes.POINTER(Opt), ctypes.c_float, _P, _P, _P, _P, _P, _I32, _P, _I64, _P,
                                       _P]),
    "dips_eval": (ctypes.c_int, [ctypes.POINTER(Net), _P, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "dips_pack_tiles": (_I64, [_P, _I64, _P, _P, _I32, _P, _I64, _P]),
    "dips_tile_slots": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), _P, _I64, _P, _P, _P, _P, _P,
                                       _P, _P, _P]),
    "dips_tile_supported": (ctypes.c_int, [ctypes.POINTER(Net)]),
    "dips_tile_partial_size": (_I64, [ctypes.POINTER(Net)]),
    "dips_tile_loss_grad": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(Tiles), _I64,
                                           _I64, ctypes.POINTER(Net), _P, _P, _I64, ctypes.c_int, ctypes.POINTER(_I32),
                                           _P]),
    "dips_tile_train_step": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(Tiles),
                                            _I64, _I64, ctypes.POINTER(Net), ctypes.POINTER(Opt), ctypes.c_float, _P,
                                            _P, _P, _P, _P, _I32, _P, _I64, _P, _P]),
    "dips_tile_train_level": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(Tiles),
                                             ctypes.POINTER(_I64), _I32, _I64, _I32, ctypes.POINTER(Net),
                                             ctypes.POINTER(Opt), _P, _P, _P, _P, _P, _I32, _P, _I64, _P, _P]),
    "dips_batch_workspace": (_I64, [ctypes.POINTER(BatchProblem), _I32, ctypes.POINTER(Net)]),
    "dips_batch_train": (ctypes.c_int, [ctypes.POINTER(BatchProblem), _I32, _I32, ctypes.POINTER(Net),
                                        ctypes.POINTER(Opt), _P, _I64, _P]),
    "dips_batch_workspace_cfg": (_I64, [ctypes.POINTER(BatchProblem), _I32, ctypes.POINTER(Net),
                                        ctypes.POINTER(BatchCfg)]),
    "dips_batch_setup": (ctypes.c_int, [ctypes.POINTER(BatchProblem), _I32, _I32, ctypes.POINTER(Net),
                                        ctypes.POINTER(Opt), ctypes.POINTER(BatchCfg), _P, _I64,
                                        ctypes.POINTER(BatchPlan), _P]),
    "dips_batch_steps": (ctypes.c_int, [ctypes.POINTER(Net), ctypes.POINTER(BatchPlan), _I32, _P, _P]),
    "dips_mfma_supported": (ctypes.c_int, [ctypes.POINTER(Net)]),
    "dips_mfma_workspace": (_I64, [ctypes.POINTER(Net), _I64, _I64]),
    "dips_mfma_train_step": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(MfmaBatch),
                                            ctypes.POINTER(Net), ctypes.POINTER(Opt), ctypes.c_float, _P, _P, _P, _P,
                                            _P, _I32, _P, _I64, _P, _I64, _P]),
    "dips_mfma_train_step_noalloc": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records),
                                                    ctypes.POINTER(MfmaBatch), ctypes.POINTER(Net), ctypes.POINTER(Opt),
                                                    ctypes.c_float, _P, _P, _P, _P, _P, _I32, _P]),
    "dips_mfma_loss_grad": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(MfmaBatch),
                                           ctypes.POINTER(Net), ctypes.c_float, _P, _P, _P, _P, _I64, _P, _I64, _P]),
    "dips_mfma_items": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(MfmaBatch), _P, _P,
                                       _P]),
    "dips_hash_table_rows": (_I64, [ctypes.POINTER(HashGrid)]),
    "dips_hash_param_count": (_I64, [ctypes.POINTER(HashGrid)]),
    "dips_hash_workspace": (_I64, [ctypes.POINTER(HashGrid), _I64, _I64, ctypes.POINTER(_I64)]),
    "dips_hash_train_step": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(MfmaBatch),
                                            ctypes.POINTER(HashGrid), ctypes.POINTER(Opt), ctypes.c_float, _P, _P,
                                            _P, _P, _P, _I32, _P, _I64, _P, _I64, _P]),
    "dips_hash_loss_grad": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), ctypes.POINTER(MfmaBatch),
                                           ctypes.POINTER(HashGrid), ctypes.c_float, _P, _P, _P, _P, _I64, _P, _I64,
                                           _P]),
    "dips_hash_eval": (ctypes.c_int, [ctypes.POINTER(HashGrid), _P, _P, _P, _I64, _P, _P, _P]),
    "dips_grid_apply": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), _P, _P, ctypes.c_int, _P]),
    "dips_bicgstab": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.POINTER(Records), _P, _P, _P, _I32, ctypes.c_double,
                                     ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_double), _P]),
    "dips_atoms_eval": (ctypes.c_int, [_P, _I32, _P, _P, _I32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _P, _I64, _P, _P, _P]),
    "dips_geom_regions": (ctypes.c_int, [_P, _I64, _P, _P]),
    "dips_geom_cells": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _I64,
                                       _P, _P, _P, _P, _P, _P]),
    "dips_geom_regression": (ctypes.c_int, [ctypes.POINTER(Grid), ctypes.c_float, ctypes.c_float, ctypes.c_float, _P,
                                            _P, _I64, _P, _P, ctypes.c_double, _P, _P, _P, _P, _P, _P]),
    "dips_geom_assemble": (ctypes.c_int, [ctypes.POINTER(Assemble), _P]),
}

_lib = None


def lib():
    """Load (once) and return the library, raising DipsError if unavailable."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DipsError(f"{LIB_PATH} is not built; run jax-dips_amd/build.py (or __graft_entry__.build())")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise DipsError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_roctx = None


class trace_range:
    """A roctx range around a block of host code (the training loop's levels and epochs); the library's
    entry points carry their own.  A no-op unless a tracer (rocprofv3 --marker-trace) is attached, or
    where the roctx runtime is absent."""

    def __init__(self, name):
        self.name = name.encode()

    def __enter__(self):
        global _roctx
        if _roctx is None:
            try:
                _roctx = ctypes.CDLL("librocprofiler-sdk-roctx.so.1")
                _roctx.roctxRangePushA.argtypes = [ctypes.c_char_p]
            except OSError:
                _roctx = False
        if _roctx:
            _roctx.roctxRangePushA(self.name)
        return self

    def __exit__(self, *exc):
        if _roctx:
            _roctx.roctxRangePop()
        return False


def check(status, what):
    if status != DIPS_OK:
        msg = lib().dips_last_error().decode()
        raise DipsError(f"{what} failed (status {status}): {msg}")


def i64_array(values):
    vals = [int(v) for v in values]
    return (ctypes.c_int64 * len(vals))(*vals)


def ptr(t):
    """Raw device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def current_stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
