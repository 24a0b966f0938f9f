"""ctypes binding of libmodle_gpu.so (include/modle_gpu.h).

The library is built in-tree by modlefork_amd.build; loading fails loudly when it is missing:
there is no CPU fallback for the simulation entry points.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmodle_gpu.so")

U64P = C.POINTER(C.c_uint64)
U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)
DP = C.POINTER(C.c_double)

MGPU_OK = 0
MGPU_EINVAL = -1
MGPU_EDEVICE = -2
MGPU_ENOMEM = -3
MGPU_ERANGE = -4
MODE_DETERMINISTIC = 0
MODE_FAST = 1
STOP_CONTACT_DENSITY = 0
STOP_EPOCHS = 1
SET_REV_SPEED = 1 << 0
SET_FWD_SPEED = 1 << 1
SET_BARRIER_OCCUPANCY = 1 << 2


class MgpuConfig(C.Structure):
    _fields_ = [
        ("bin_size", C.c_uint64),
        ("diagonal_width", C.c_uint64),
        ("sample_loop_contacts", C.c_int32),
        ("sample_tad_contacts", C.c_int32),
        ("noisify_contacts", C.c_int32),
        ("stopping_criterion", C.c_int32),
        ("tad_to_loop_contact_ratio", C.c_double),
        ("genextreme_mu", C.c_double),
        ("genextreme_sigma", C.c_double),
        ("genextreme_xi", C.c_double),
        ("rev_extrusion_speed", C.c_uint64),
        ("fwd_extrusion_speed", C.c_uint64),
        ("rev_extrusion_speed_std", C.c_double),
        ("fwd_extrusion_speed_std", C.c_double),
        ("number_of_lefs_per_mbp", C.c_double),
        ("avg_lef_processivity", C.c_uint64),
        ("contact_sampling_interval", C.c_uint64),
        ("extrusion_barrier_occupancy", C.c_double),
        ("barrier_occupied_stp", C.c_double),
        ("barrier_not_occupied_stp", C.c_double),
        ("hard_stall_lef_stability_multiplier", C.c_double),
        ("soft_stall_lef_stability_multiplier", C.c_double),
        ("probability_of_extrusion_unit_bypass", C.c_double),
        ("lef_bar_major_collision_pblock", C.c_double),
        ("lef_bar_minor_collision_pblock", C.c_double),
        ("target_simulation_epochs", C.c_uint64),
        ("target_contact_density", C.c_double),
        ("num_cells", C.c_uint64),
        ("seed", C.c_uint64),
        ("probability_normalization_factor", C.c_uint64),
        ("normalize_probabilities", C.c_int32),
        ("skip_burnin", C.c_int32),
        ("burnin_history_length", C.c_uint64),
        ("burnin_smoothing_window_size", C.c_uint64),
        ("min_burnin_epochs", C.c_uint64),
        ("max_burnin_epochs", C.c_uint64),
        ("burnin_speed_coefficient", C.c_double),
        ("track_1d_lef_position", C.c_int32),
        ("log_model_internal_state", C.c_int32),
    ]


class MgpuParams(C.Structure):
    _fields_ = [
        ("bin_size", C.c_uint64),
        ("diagonal_width", C.c_uint64),
        ("rev_speed", C.c_uint64),
        ("fwd_speed", C.c_uint64),
        ("rev_speed_burnin", C.c_uint64),
        ("fwd_speed_burnin", C.c_uint64),
        ("rev_speed_std", C.c_double),
        ("fwd_speed_std", C.c_double),
        ("prob_release", C.c_double),
        ("prob_release_burnin", C.c_double),
        ("p_bypass", C.c_double),
        ("pblock_major", C.c_double),
        ("pblock_minor", C.c_double),
        ("hard_stall_mult", C.c_double),
        ("soft_stall_mult", C.c_double),
        ("tad_to_loop_ratio", C.c_double),
        ("gev_mu", C.c_double),
        ("gev_sigma", C.c_double),
        ("gev_xi", C.c_double),
        ("noisify", C.c_int32),
        ("skip_burnin", C.c_int32),
        ("contact_sampling_interval", C.c_uint64),
        ("burnin_history_length", C.c_uint64),
        ("burnin_smoothing_window", C.c_uint64),
        ("min_burnin_epochs", C.c_uint64),
        ("max_burnin_epochs", C.c_uint64),
        ("burnin_target_epochs_for_lef_activation", C.c_uint64),
        ("target_contact_density", C.c_double),
        ("target_simulation_epochs", C.c_uint64),
        ("seed", C.c_uint64),
        ("lefs_per_mbp", C.c_double),
        ("barrier_occupied_stp", C.c_double),
        ("barrier_not_occupied_stp", C.c_double),
        ("override_barrier_occupancy", C.c_int32),
        ("track_1d_lef_position", C.c_int32),
        ("log_model_internal_state", C.c_int32),
        ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class MgpuChrom(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("name_len", C.c_uint64),
        ("size", C.c_uint64),
        ("start", C.c_uint64),
        ("end", C.c_uint64),
    ]


class MgpuBarriers(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("pos", U64P),
        ("dir", U8P),
        ("stp_active", DP),
        ("stp_inactive", DP),
    ]


class MgpuBedRecord(C.Structure):
    _fields_ = [
        ("chrom_start", C.c_uint64),
        ("chrom_end", C.c_uint64),
        ("score", C.c_double),
        ("strand", C.c_char),
    ]


class MgpuWindowTask(C.Structure):
    _fields_ = [
        ("chrom", MgpuChrom),
        ("barriers", MgpuBarriers),
        ("nlefs", C.c_uint64),
        ("cell_id", C.c_uint64),
        ("target_contacts", C.c_uint64),
        ("target_epochs", C.c_uint64),
    ]


class MgpuStats(C.Structure):
    _fields_ = [
        ("cells", C.c_uint64),
        ("epochs", C.c_uint64),
        ("burnin_epochs", C.c_uint64),
        ("contacts", C.c_uint64),
        ("missed", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("dropped", C.c_uint64),
        ("dropped_occupancy", C.c_uint64),
    ]


class MgpuTuning(C.Structure):
    _fields_ = [
        ("threads_per_cell", C.c_uint32),
        ("grid", C.c_uint32),
        ("global_working_set", C.c_uint32),
        ("launches", C.c_uint32),
        ("tier_threads", C.c_uint32 * 5),
        ("tier_grid", C.c_uint32 * 5),
    ]


class MgpuLaunchStats(C.Structure):
    _fields_ = [
        ("threads", C.c_uint32),
        ("grid", C.c_uint32),
        ("cells", C.c_uint64),
        ("epochs", C.c_uint64),
        ("contacts", C.c_uint64),
        ("barriers", C.c_uint64),
        ("ms", C.c_double),
    ]


class MgpuStateRecord(C.Structure):
    _fields_ = [
        ("chrom", C.c_uint64),
        ("cell_id", C.c_uint64),
        ("epoch", C.c_uint64),
        ("burnin", C.c_int32),
        ("num_active_lefs", C.c_uint32),
        ("num_barriers", C.c_uint64),
        ("active_barriers", C.c_uint64),
        ("stalls_rev", C.c_uint64),
        ("stalls_fwd", C.c_uint64),
        ("stalls_both", C.c_uint64),
        ("lef_bar", C.c_uint64),
        ("primary", C.c_uint64),
        ("secondary", C.c_uint64),
        ("avg_loop_size", C.c_double),
    ]


class MgpuSgConfig(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "in_len", "out_offset", "out_len", "ndiag", "in_ch", "enc_ch", "enc_layers", "enc_k",
        "head_ch", "head_layers", "max_batch", "reserved")] + [("seed", C.c_uint64),
                                                                ("enc_dilation", C.c_uint32 * 16)]


FP = C.POINTER(C.c_float)

PHASE_NAMES = ["stop test and burn-in", "bind", "rank", "rank fwd (batched)", "contacts", "moves",
               "adjust", "barrier states and ends", "LEF-BAR", "primary LEF-LEF",
               "move corrections", "secondary (both sides)", "unused", "bypass fix-ups",
               "extrude and release", "unused"]
# the wave-parallel deterministic kernel's phases (mgpu_debug_phase_times entries 16-27)
DET_PHASE_NAMES = ["burn-in", "bind", "rank", "contacts", "moves and adjust", "barrier states",
                   "chromosome ends", "LEF-BAR", "primary and corrections", "secondary",
                   "bypass fix-ups", "extrude and release"]

# (name, restype, argtypes) for every entry point of include/modle_gpu.h
SIGNATURES = [
    ("mgpu_init", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("mgpu_destroy", None, [C.c_void_p]),
    ("mgpu_last_error", C.c_char_p, [C.c_void_p]),
    ("mgpu_version", C.c_char_p, []),
    ("mgpu_config_defaults", C.c_int, [C.POINTER(MgpuConfig)]),
    ("mgpu_config_transform", C.c_int, [C.POINTER(MgpuConfig), C.c_uint32, C.POINTER(MgpuParams)]),
    ("mgpu_stp_active_from_occupancy", C.c_double, [C.c_double, C.c_double]),
    ("mgpu_occupancy_from_stp", C.c_double, [C.c_double, C.c_double]),
    ("mgpu_prepare_barriers", C.c_int, [C.POINTER(MgpuParams), C.POINTER(MgpuBedRecord), C.c_uint64,
                                        C.c_uint64, C.c_uint64, U64P, U8P, DP, DP, U64P]),
    ("mgpu_import_barriers", C.c_int, [C.POINTER(MgpuParams), C.POINTER(MgpuBedRecord), DP,
                                       C.c_uint64, C.c_uint64, C.c_uint64, U64P, U64P, U64P, U8P,
                                       DP, DP, U64P]),
    ("mgpu_sort_barriers", C.c_int, [C.c_uint64, U64P, U8P, DP, DP]),
    ("mgpu_pdq_order", C.c_int, [U64P, C.c_uint64, U32P]),
    ("mgpu_apply_deletion", C.c_int, [U64P, C.c_uint64, C.c_uint64, C.c_uint64, DP, DP, U64P]),
    ("mgpu_prepare_barriers_name_puu", C.c_int, [C.POINTER(MgpuParams), C.POINTER(MgpuBedRecord),
                                                 DP, C.c_uint64, C.c_uint64, C.c_uint64, U64P, U8P,
                                                 DP, DP, U64P]),
    ("mgpu_plan_chrom", C.c_int, [C.POINTER(MgpuParams), C.POINTER(MgpuChrom), C.c_uint64, U64P,
                                  U64P, U64P, U64P, U64P]),
    ("mgpu_cell_seed", C.c_uint64, [C.POINTER(MgpuChrom), C.c_uint64, C.c_uint64]),
    ("mgpu_simulate_chrom", C.c_int, [C.c_void_p, C.POINTER(MgpuParams), C.POINTER(MgpuChrom),
                                      C.POINTER(MgpuBarriers), C.c_uint64, C.c_uint64, C.c_uint64,
                                      U64P, C.c_uint64, C.c_int, U32P, U64P,
                                      C.POINTER(MgpuStats)]),
    ("mgpu_job_create", C.c_int, [C.c_void_p, C.POINTER(MgpuParams), C.c_uint64,
                                  C.POINTER(MgpuChrom), C.POINTER(MgpuBarriers), U64P, U64P,
                                  C.POINTER(U64P), U64P, C.c_int, C.POINTER(C.c_void_p)]),
    ("mgpu_job_create_multi", C.c_int, [C.c_void_p, C.POINTER(MgpuParams), C.c_uint64,
                                        C.POINTER(MgpuChrom), C.POINTER(MgpuBarriers), U64P, U64P,
                                        C.POINTER(U64P), U64P, C.c_int, C.POINTER(C.c_void_p)]),
    ("mgpu_job_create_windows", C.c_int, [C.c_void_p, C.POINTER(MgpuParams), C.c_uint64,
                                          C.POINTER(MgpuWindowTask), C.c_int,
                                          C.POINTER(C.c_void_p)]),
    ("mgpu_job_block_sums", C.c_int, [C.c_void_p, C.c_uint64, U64P, U64P, U64P, C.c_uint64,
                                      U32P]),
    ("mgpu_band_block_sums", C.c_int, [C.c_void_p, U32P, C.c_uint64, C.c_uint64, C.c_uint64,
                                       U64P, U64P, C.c_uint64, U32P]),
    ("mgpu_job_run", C.c_int, [C.c_void_p, C.POINTER(MgpuStats)]),
    ("mgpu_job_capture", C.c_int, [C.c_void_p]),
    ("mgpu_job_release_graph", C.c_int, [C.c_void_p]),
    ("mgpu_job_read_band", C.c_int, [C.c_void_p, C.c_uint64, U32P, U64P]),
    ("mgpu_job_band_device_ptr", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p), U64P]),
    ("mgpu_job_destroy", None, [C.c_void_p]),
    ("mgpu_job_launch_stats", C.c_int, [C.c_void_p, C.POINTER(MgpuLaunchStats), C.c_uint32,
                                        C.POINTER(C.c_uint32)]),
    ("mgpu_job_read_occupancy", C.c_int, [C.c_void_p, C.c_uint64, U64P, U64P]),
    ("mgpu_job_set_state_log", C.c_int, [C.c_void_p, C.c_uint64]),
    ("mgpu_job_read_state_log", C.c_int, [C.c_void_p, C.POINTER(MgpuStateRecord), C.c_uint64,
                                          U64P]),
    ("mgpu_job_set_tuning", C.c_int, [C.c_void_p, C.POINTER(MgpuTuning)]),
    ("mgpu_job_get_tuning", C.c_int, [C.c_void_p, C.POINTER(MgpuTuning)]),
    ("mgpu_job_autotune", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(MgpuTuning),
                                    C.POINTER(C.c_float)]),
    ("mgpu_debug_step", C.c_int, [C.c_void_p, C.POINTER(MgpuParams), C.POINTER(MgpuChrom), C.c_int,
                                  C.c_uint64, U64P, U64P, U64P, U64P, U64P, U64P, U64P, U64P, U64P,
                                  C.POINTER(MgpuBarriers), U8P, U64P]),
    ("mgpu_debug_draws", C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_uint64, U64P]),
    ("mgpu_debug_philox", C.c_int, [C.c_void_p, U32P, C.c_uint64, U32P]),
    ("mgpu_simulate_chrom_noalloc", C.c_int, [C.c_void_p, C.POINTER(MgpuParams),
                                              C.POINTER(MgpuChrom), C.POINTER(MgpuBarriers),
                                              C.c_uint64, C.c_uint64, C.c_uint64, U64P, C.c_uint64,
                                              C.c_void_p, U64P, C.POINTER(MgpuStats)]),
    ("mgpu_debug_fast_math", C.c_int, [C.c_void_p, C.c_int, DP, C.c_uint64, DP]),
    ("mgpu_debug_band_increments", C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64,
                                             C.c_uint64, U64P, U64P, U32P, U64P]),
    ("mgpu_debug_collisions", C.c_int, [C.c_void_p, C.c_uint64, U64P, U8P, U64P, U64P, U32P,
                                        U32P]),
    ("mgpu_debug_phase_times", C.c_int, [C.c_void_p, C.c_int, U64P, C.c_uint64]),
    ("mgpu_debug_fast_serial_fixups", C.c_int, [C.c_void_p, C.c_int]),
    ("mgpu_debug_det_kernel", C.c_int, [C.c_void_p, C.c_int]),
    ("mgpu_job_kernel", C.c_int, [C.c_void_p, C.c_char_p, C.c_uint64]),
    ("mgpu_debug_det_slab_slots", C.c_int, [C.c_void_p, C.c_uint64]),
    ("mgpu_job_cell_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64,
                                      C.POINTER(C.c_uint64)]),
    ("mgpu_context_stream", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]),
    ("mgpu_sg_config_default", C.c_int, [C.POINTER(MgpuSgConfig)]),
    ("mgpu_sg_create", C.c_int, [C.c_void_p, C.POINTER(MgpuSgConfig), C.POINTER(C.c_void_p)]),
    ("mgpu_sg_destroy", None, [C.c_void_p]),
    ("mgpu_sg_last_error", C.c_char_p, [C.c_void_p]),
    ("mgpu_sg_num_params", C.c_uint64, [C.c_void_p]),
    ("mgpu_sg_get_params", C.c_int, [C.c_void_p, FP]),
    ("mgpu_sg_set_params", C.c_int, [C.c_void_p, FP]),
    ("mgpu_sg_forward", C.c_int, [C.c_void_p, FP, C.c_uint64, FP]),
    ("mgpu_sg_gradients", C.c_int, [C.c_void_p, FP, FP, C.c_uint64, FP, FP]),
    ("mgpu_sg_adamw_step", C.c_int, [C.c_void_p, FP, C.c_float, C.c_float]),
    ("mgpu_sg_set_data", C.c_int, [C.c_void_p, FP, FP, C.c_uint64]),
    ("mgpu_sg_train_step", C.c_int, [C.c_void_p, U32P, C.c_uint64, C.c_float, C.c_float, FP, FP]),
    ("mgpu_sg_predict_timed", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, FP]),
    ("mgpu_sg_reserve_bands", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    ("mgpu_sg_predict_chromosomes", C.c_int, [C.c_void_p, FP, C.c_uint64, C.c_uint64, C.c_uint64,
                                              U32P]),
    ("mgpu_sg_read_bands", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, FP]),
    ("mgpu_debug_mfma", C.c_int, [C.c_void_p, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), FP]),
    ("mgpu_debug_cell_trace", C.c_int, [C.c_void_p, C.POINTER(MgpuParams), C.POINTER(MgpuChrom),
                                        C.POINTER(MgpuBarriers), C.c_uint64, C.c_uint64, C.c_uint64,
                                        C.c_uint64, C.c_uint64, U64P, U64P]),
]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Loads the in-tree library; raises LibraryMissing when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -m modlefork_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class MgpuError(RuntimeError):
    pass


def check(rc: int, ctx=None):
    if rc != 0:
        msg = lib().mgpu_last_error(ctx)
        raise MgpuError(f"libmodle_gpu error {rc}: {msg.decode() if msg else ''}")
