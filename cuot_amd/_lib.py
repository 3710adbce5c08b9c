"""ctypes loader for libcuot_amd.so (the gfx950 engine) and its C ABI.

The library is built in-tree by ``python -m cuot_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly if it is missing: there is
no CPU fallback behind this interface.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcuot_amd.so")

P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
INT = C.c_int
U64 = C.c_uint64


class Block(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


class GgmParams(C.Structure):
    _fields_ = [("t", I64), ("h", I32), ("mode", I32), ("n_out", I64), ("k0", Block), ("k1", Block)]


class LpnParams(C.Structure):
    _fields_ = [("n", I64), ("k", I64), ("d", I32), ("ntasks", I32), ("seed", Block)]


class SotConfig(C.Structure):
    _fields_ = [("n", I64), ("t", I64), ("h", I32), ("reserved", I32), ("k0", Block), ("k1", Block),
                ("seed", Block)]


class FerretConfig(C.Structure):
    _fields_ = [("n", I64), ("k", I64), ("t", I64), ("h", I32), ("ntasks", I32), ("seed", Block), ("n_pre", I64)]


class SilentConfig(C.Structure):
    _fields_ = [("id", I32), ("logOT", I32), ("nTree", C.c_uint64), ("leftKey", C.c_uint32 * 4),
                ("rightKey", C.c_uint32 * 4), ("gpuPerParty", I32), ("reserved", I32)]


class Instance(C.Structure):
    _fields_ = [("kind", I32), ("h", I32), ("n", I64), ("t", I64), ("k", I64), ("ntasks", I32), ("client", I32),
                ("k0", Block), ("k1", Block), ("seed", Block)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_batch", I32), ("streams", I32), ("qc_chunk_bytes", C.c_size_t),
                ("use_graph", I32), ("plain", I32), ("max_batch_cots", I64), ("arena_d", P),
                ("ferret_clients", I32), ("reserved2", I32), ("ferret_session", FerretConfig)]


class ServerResult(C.Structure):
    _fields_ = [("done", I32), ("status", I32), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double), ("out0", P), ("out1", P), ("bits", P), ("delta", P)]


class KnobValues(C.Structure):
    _fields_ = [("qc_radix", I32), ("qc_cols", I32), ("ggm_dfs", I32), ("lpn_wgs", I32), ("qc_mfma_max", I32),
                ("aes_impl", I32)]


class LaunchTime(C.Structure):
    _fields_ = [("kernel", C.c_char * 32), ("ms", C.c_float), ("bytes", C.c_double)]


# name -> (restype, argtypes); every symbol declared in include/cuot_amd.h
SIGNATURES = {
    "cuot_init": (INT, [INT, C.POINTER(P)]),
    "cuot_destroy": (INT, [P]),
    "cuot_last_error": (C.c_char_p, []),
    "cuot_device_sync": (INT, [P]),
    "cuot_aes_ecb": (INT, [P, C.POINTER(Block), P, P, I64, P]),
    "cuot_prg": (INT, [P, C.POINTER(Block), U64, U64, I64, P, P]),
    "cuot_ggm_sender": (INT, [P, C.POINTER(GgmParams), P, P, P, P]),
    "cuot_ggm_sender_bs_work_bytes": (C.c_size_t, [C.POINTER(GgmParams)]),
    "cuot_ggm_sender_bs": (INT, [P, C.POINTER(GgmParams), P, P, P, P, C.c_size_t, P]),
    "cuot_ggm_recver_bs_work_bytes": (C.c_size_t, [C.POINTER(GgmParams)]),
    "cuot_ggm_recver_bs": (INT, [P, C.POINTER(GgmParams), P, P, P, P, P, C.c_size_t, P]),
    "cuot_ggm_recver": (INT, [P, C.POINTER(GgmParams), P, P, P, P, P]),
    "cuot_lpn": (INT, [P, C.POINTER(LpnParams), P, P, P, P]),
    "cuot_lpn_dual": (INT, [P, C.POINTER(LpnParams), P, P, P, P, P, P, P]),
    "cuot_qc_ahat_len": (I64, [I64]),
    "cuot_qc_scratch_bytes": (C.c_size_t, [I64]),
    "cuot_qc_prepare": (INT, [P, I64, P, P, P, P]),
    "cuot_qc_encode": (INT, [P, I64, P, P, P, P, P]),
    "cuot_qc_sparse": (INT, [P, I64, P, P, I64, P, P]),
    "cuot_bit_transpose": (INT, [P, P, I64, P, P]),
    "cuot_bit_untranspose": (INT, [P, P, I64, P, P]),
    "cuot_sot_workspace_bytes": (C.c_size_t, [C.POINTER(SotConfig)]),
    "cuot_sot_run": (INT, [P, C.POINTER(SotConfig), P, C.c_size_t, P, P, P, P, P]),
    "cuot_sot_run_stages": (INT, [P, C.POINTER(SotConfig), P, C.c_size_t, P, P, P, P, C.c_uint32, P]),
    "cuot_ferret_workspace_bytes": (C.c_size_t, [C.POINTER(FerretConfig)]),
    "cuot_ferret_run_stages": (INT, [P, C.POINTER(FerretConfig), P, C.c_size_t, P, P, P, C.c_uint32, P]),
    "cuot_setdev": (INT, [INT]),
    "cuot_mpcot_sender": (INT, [P, P, P, P, P, P, P, P, INT, P, INT]),
    "cuot_mpcot_recver": (INT, [P, P, P, P, P, P, P, INT, P, INT]),
    "cuot_primal_lpn": (INT, [P, INT, P, I64, P, INT, P, P, P, P, P, INT]),
    "cuot_sot_sender_new": (INT, [P, C.POINTER(SilentConfig), C.POINTER(Block), C.POINTER(P)]),
    "cuot_sot_recver_new": (INT, [P, C.POINTER(SilentConfig), C.POINTER(Block), C.POINTER(P)]),
    "cuot_sot_role_free": (INT, [P]),
    "cuot_sot_base_ot": (INT, [P]),
    "cuot_sot_seed_expand": (INT, [P]),
    "cuot_sot_get_punc_key": (INT, [P, P]),
    "cuot_sot_dual_lpn": (INT, [P]),
    "cuot_sot_outputs": (INT, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(I64)]),
    "cuot_ferret_run": (INT, [P, C.POINTER(FerretConfig), P, C.c_size_t, P, P, P, P]),
    "cuot_ferret_run_noalloc": (INT, [P, C.POINTER(FerretConfig), P, P, P, P]),
    "cuot_batch_new": (INT, [P, C.POINTER(Instance), INT, C.c_size_t, C.POINTER(P)]),
    "cuot_batch_workspace_bytes": (C.c_size_t, [P]),
    "cuot_batch_bind": (INT, [P, P, C.c_size_t, C.POINTER(P), C.POINTER(P), C.POINTER(P), P]),
    "cuot_batch_run": (INT, [P, C.c_uint32, P]),
    "cuot_batch_free": (INT, [P]),
    "cuot_server_new": (INT, [P, C.POINTER(ServerConfig), C.POINTER(P)]),
    "cuot_server_submit": (INT, [P, C.POINTER(Instance), C.POINTER(I64)]),
    "cuot_server_poll": (INT, [P, I64, INT, C.POINTER(ServerResult)]),
    "cuot_server_release": (INT, [P, I64]),
    "cuot_server_stats": (INT, [P, C.POINTER(I64), C.POINTER(C.c_size_t)]),
    "cuot_server_free": (INT, [P]),
    "cuot_mlp_param_count": (INT, [INT, INT]),
    "cuot_mlp_workspace_bytes": (C.c_size_t, [INT, INT, I64]),
    "cuot_mlp_forward": (INT, [P, P, INT, INT, P, I64, P, P]),
    "cuot_mlp_grad": (INT, [P, P, INT, INT, P, P, I64, P, C.c_size_t, P, P, P]),
    "cuot_mlp_adam": (INT, [P, P, P, P, P, INT, INT, C.c_float, P]),
    "cuot_set_knobs": (INT, [C.POINTER(KnobValues)]),
    "cuot_get_knobs": (INT, [C.POINTER(KnobValues)]),
    "cuot_qc_mfma_crossover": (I64, []),
    "cuot_timing_begin": (INT, [P]),
    "cuot_timing_end": (INT, [P, C.POINTER(LaunchTime), INT, C.POINTER(INT)]),
    "cuot_ferret_session_new": (INT, [P, C.POINTER(FerretConfig), C.POINTER(P)]),
    "cuot_ferret_session_info": (INT, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)]),
    "cuot_ferret_session_delta": (INT, [P, P, P]),
    "cuot_ferret_session_extend": (INT, [P, P, P, P]),
    "cuot_ferret_session_free": (INT, [P]),
    "cuot_ferret_session_extend_stages": (INT, [P, P, P, INT, P]),
    "cuot_ferret_session_set_malicious": (INT, [P, INT]),
    "cuot_cot_online": (INT, [P, P, P, P, P, P, I64, P]),
    "cuot_cot_to_ot": (INT, [P, P, P, P, Block, C.c_uint64, P, P, P, P, P, I64, P]),
    "cuot_cot_to_rot": (INT, [P, P, P, P, Block, C.c_uint64, P, P, P, I64, P]),
    "cuot_ferret_session_check": (INT, [P, C.POINTER(INT), P, P, P]),
    "cuot_ferret_state_size": (I64, [C.POINTER(FerretConfig)]),
    "cuot_ferret_session_assemble_state": (INT, [P, INT, P, I64]),
    "cuot_ferret_session_disassemble_state": (INT, [P, C.POINTER(FerretConfig), I64, P, P, I64, C.POINTER(P)]),
    "cuot_ferret_session_write_pre_file": (INT, [P, INT, C.c_char_p]),
    "cuot_ferret_session_read_pre_files": (INT, [P, C.POINTER(FerretConfig), I64, C.c_char_p, C.c_char_p,
                                                 C.POINTER(P)]),
    "cuot_serve_replay": (INT, [P, C.POINTER(ServerConfig), C.POINTER(Instance), C.POINTER(C.c_double), INT,
                                C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(I64), C.POINTER(I64)]),
}

_lib = None


class CuotError(RuntimeError):
    pass


def load():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CuotError(f"{LIB_PATH} not found: build it with `python -m cuot_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().cuot_last_error().decode(errors="replace")
        raise CuotError(f"{what} failed with status {rc}: {msg}")
