"""ctypes binding of ``libbetse_amd.so`` (declarations in ``include/betse_amd.h``).

The library is built in-tree by ``betse_amd.build.build_library`` (``hipcc
--offload-arch=gfx950``). Loading fails loudly when it is missing: there is no
CPU fallback for the step.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

LIB_NAME = "libbetse_amd.so"
# the in-tree build; BETSE_AMD_LIBRARY names another build of the same sources
LIB_PATH = os.environ.get("BETSE_AMD_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

BT_OK = 0
BT_UNSTABLE = 1
BT_EINVAL = -1
BT_EHIP = -2
BT_MAX_IONS = 8
BT_MAX_CHAN = 4
BT_CHAN_KINDS = {"Nav1p2": 0, "Kv1p1": 1, "Cav1p2": 2}
BT_FLUX_FP64 = 0
BT_FLUX_FP32 = 1

BT_TILE_MEMBRANES = 64
BT_TILE_CELLS = 32

# bt_time_kernel ids (env_field_exchange: the env kernel of exchange_form 2, which does the exchange)
KERNEL_IDS = {"mem_cell": 0, "env_field": 1, "env_exchange": 2, "env_field_only": 3, "network": 4,
              "env_field_exchange": 1}

BT_NET_MAX_SPECIES = 16
BT_NET_MAX_RATES = 32
BT_NET_MAX_HILL = 4
BT_NET_MAX_TERMS = 4
BT_NET_MAX_STOICH = 8
BT_RATE_GROWTH = 0
BT_RATE_REACTION = 1
BT_GNN_SCRATCH_FLOATS = 1 << 18

_d8 = ctypes.c_double * BT_MAX_IONS
_d5 = ctypes.c_double * 5
_i4 = ctypes.c_int32 * BT_MAX_CHAN
_d4 = ctypes.c_double * BT_MAX_CHAN
_vp = ctypes.c_void_p


class BtRate(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("species", ctypes.c_int32), ("reversible", ctypes.c_int32),
        ("n_act", ctypes.c_int32), ("n_inh", ctypes.c_int32), ("n_reac", ctypes.c_int32),
        ("n_prod", ctypes.c_int32), ("n_stoich", ctypes.c_int32),
        ("act", ctypes.c_int32 * BT_NET_MAX_HILL), ("inh", ctypes.c_int32 * BT_NET_MAX_HILL),
        ("reac", ctypes.c_int32 * BT_NET_MAX_TERMS), ("prod", ctypes.c_int32 * BT_NET_MAX_TERMS),
        ("stoich_species", ctypes.c_int32 * BT_NET_MAX_STOICH),
        ("act_km", ctypes.c_double * BT_NET_MAX_HILL), ("act_n", ctypes.c_double * BT_NET_MAX_HILL),
        ("inh_km", ctypes.c_double * BT_NET_MAX_HILL), ("inh_n", ctypes.c_double * BT_NET_MAX_HILL),
        ("reac_coeff", ctypes.c_double * BT_NET_MAX_TERMS), ("reac_km", ctypes.c_double * BT_NET_MAX_TERMS),
        ("prod_coeff", ctypes.c_double * BT_NET_MAX_TERMS), ("prod_km", ctypes.c_double * BT_NET_MAX_TERMS),
        ("stoich_coeff", ctypes.c_double * BT_NET_MAX_STOICH),
        ("r_prod", ctypes.c_double), ("r_decay", ctypes.c_double), ("decay_max", ctypes.c_double),
        ("gad_scale", ctypes.c_double), ("vmax", ctypes.c_double), ("keq", ctypes.c_double),
    ]


class BtNetwork(ctypes.Structure):
    _fields_ = [("n_species", ctypes.c_int32), ("n_rates", ctypes.c_int32),
                ("rates", BtRate * BT_NET_MAX_RATES)]


class BtCluster(ctypes.Structure):
    _fields_ = [
        ("grid_ny", ctypes.c_int32), ("grid_nx", ctypes.c_int32), ("sq_off", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("delta", ctypes.c_double), ("mean_sa", ctypes.c_double), ("ko_env", ctypes.c_double),
        ("alpha_NaK", ctypes.c_double),
    ]


class BtParams(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int32), ("n_mems", ctypes.c_int32), ("n_squares", ctypes.c_int32),
        ("reserved0", ctypes.c_int32), ("n_ions", ctypes.c_int32),
        ("i_na", ctypes.c_int32), ("i_k", ctypes.c_int32), ("i_ca", ctypes.c_int32),
        ("flux_mode", ctypes.c_int32), ("n_chan", ctypes.c_int32),
        ("chan_kind", _i4), ("chan_ion", _i4), ("chan_max_dm", _d4), ("chan_dt", ctypes.c_double),
        ("z", _d8), ("d_gj", _d8), ("c_bound", _d8),
        ("dt", ctypes.c_double), ("T", ctypes.c_double), ("T_gj", ctypes.c_double),
        ("F", ctypes.c_double), ("R", ctypes.c_double), ("q", ctypes.c_double), ("kb", ctypes.c_double),
        ("cm", ctypes.c_double), ("tm", ctypes.c_double),
        ("gj_len", ctypes.c_double), ("gj_surface", ctypes.c_double),
        ("NaK_block", ctypes.c_double), ("deltaGATP", ctypes.c_double),
        ("cATP", ctypes.c_double), ("cADP", ctypes.c_double), ("cPi", ctypes.c_double),
        ("KmNK_Na", ctypes.c_double), ("KmNK_K", ctypes.c_double), ("KmNK_ATP", ctypes.c_double),
        ("alpha_Ca", ctypes.c_double), ("KmCa_Ca", ctypes.c_double), ("KmCa_ATP", ctypes.c_double),
        ("cell_height", ctypes.c_double),
        ("eo", ctypes.c_double), ("er", ctypes.c_double),
        ("cell_radius", ctypes.c_double), ("true_cell_size", ctypes.c_double),
        ("gauss_w", _d5),
        ("smooth_cells", ctypes.c_double), ("mu", ctypes.c_double),
    ]


class BtBuffers(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "mem_ptr", "mem_cell", "nn_mem", "nn_cell", "mem_sq", "sq_ptr", "sq_mems", "mem_pos", "tile_ptr",
        "tile_cluster", "sq_cluster",
        "mem_sa", "gjopen", "cell_vol", "diviterm", "Dm", "D_env", "sq_has_mem",
        "c_cell", "c_lag", "vm_cell", "c_env", "E_x", "E_y", "v_env", "mem_nx", "mem_ny", "cell_sa",
        "gates", "chan_mod", "Phi_b")] + [
        ("n_tiles", ctypes.c_int32), ("tile_m", ctypes.c_int32)]


OBS_FIELDS = ("fluxes_mem", "fluxes_gj", "fluxes_env_x", "fluxes_env_y", "chan_flux", "rate_NaKATP",
              "Jn", "I_mem", "E_gj_x", "E_gj_y", "J_cell_x", "J_cell_y", "vm_ave",
              "J_env_x", "J_env_y", "Jtx", "Jty", "B_field", "vm")


class BtObs(ctypes.Structure):
    _fields_ = [(n, _vp) for n in OBS_FIELDS]


class BtTuning(ctypes.Structure):
    _fields_ = [("mem_waves", ctypes.c_int32), ("env_ions_per_block", ctypes.c_int32),
                ("exchange_form", ctypes.c_int32), ("graph_pairs", ctypes.c_int32), ("persistent", ctypes.c_int32),
                ("streams", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


# bt_autotune's trial order
AUTOTUNE_TRIALS = ("mem_waves=1", "mem_waves=2", "env_ions_per_block=1", "env_ions_per_block=2",
                   "env_ions_per_block=3", "env_ions_per_block=n_ions", "exchange_form=0", "exchange_form=1",
                   "persistent=0 (ms per step)", "persistent=1 (ms per step)", "streams=2 (ms per step)",
                   "streams=4 (ms per step)", "exchange_form=2 (ms per step)")

BT_MAX_STREAMS = 4
BT_EXCHANGE_FUSED = 2


# symbol -> (restype, argtypes); every function include/betse_amd.h declares
SIGNATURES = {
    "bt_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(BtParams), ctypes.POINTER(BtCluster),
                                  ctypes.c_int32, ctypes.POINTER(BtBuffers), _vp]),
    "bt_destroy": (ctypes.c_int, [_vp]),
    "bt_validate": (ctypes.c_int, [ctypes.POINTER(BtParams), ctypes.POINTER(BtCluster), ctypes.c_int32,
                                    ctypes.POINTER(BtBuffers)]),
    "bt_capture_graphs": (ctypes.c_int, [_vp]),
    "bt_refresh_params": (ctypes.c_int, [_vp]),
    "bt_stage_boundary": (ctypes.c_int, [_vp, _vp]),
    "bt_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(BtTuning)]),
    "bt_set_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(BtTuning)]),
    "bt_autotune": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(BtTuning), ctypes.POINTER(ctypes.c_double),
                                    ctypes.c_int32]),
    "bt_update_v": (ctypes.c_int, [_vp]),
    "bt_step": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "bt_sync": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "bt_step_observed": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(BtObs)]),
    "bt_observables": (ctypes.c_int, [_vp, ctypes.POINTER(BtObs)]),
    "bt_gemm_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "bt_fp32_math": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "bt_fp64_math": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "bt_time_kernel": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "bt_set_network": (ctypes.c_int, [_vp, _vp, _vp]),
    "bt_residual": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "bt_steps_done": (ctypes.c_int64, [_vp]),
    "bt_last_error": (ctypes.c_char_p, [_vp]),
    "bt_version": (ctypes.c_int, []),
    # warm-start graph network
    "bt_gnn_gemm": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int64,
                                   _vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int32, _vp, ctypes.c_int32, _vp]),
    "bt_gnn_bias_silu": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "bt_gnn_bias": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "bt_gnn_silu_bwd": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp]),
    "bt_gnn_mean_agg": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32,
                                       ctypes.c_int32, _vp]),
    "bt_gnn_mean_agg_T": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, _vp]),
    "bt_gnn_colsum": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _vp, _vp,
                                     ctypes.c_int32, _vp]),
    "bt_gnn_colsum_noalloc": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    "bt_gnn_mse": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_float, _vp, _vp, ctypes.c_int32, _vp]),
    "bt_gnn_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_int32, _vp]),
}

_lib: Optional[ctypes.CDLL] = None


def load(path: Optional[str] = None) -> ctypes.CDLL:
    """Load the HIP library (once). Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the gfx950 library first "
            "(python -c 'import __graft_entry__ as g; g.build()'). The GPU step has no CPU fallback.")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
