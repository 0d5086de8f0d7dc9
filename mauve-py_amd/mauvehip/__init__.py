"""mauvehip -- Python host layer over libmauvehip.so (gfx950).

Mirrors the reference's interfaces for the seed-anchoring and inter-anchor DP
path:

* ``SortedMerList``      -- DNAFileSML / DNAMemorySML (LM/SortedMerList.h:69-282)
* ``find_matches``       -- PairwiseMatchFinder::FindMatches (LM/MemHash.cpp:109-127)
* ``global_align``       -- GlobalAlign / NWSmall for a batch of jobs
                            (MU/glbalign.cpp:69-81, MU/nwsmall.cpp:500-670)
* ``get_seed`` & co      -- LM/SeedMasks.h:298-401

Every call goes through the C ABI declared in include/mauvehip.h.  There is
no CPU fallback: if the library is missing or no GPU is visible, the calls
raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmauvehip.so")
MH_MAX_SEQ = 4

MH_OK = 0
ERRORS = {
    -1: "MH_ERR_INVALID",
    -2: "MH_ERR_DEVICE",
    -3: "MH_ERR_NOMEM",
    -4: "MH_ERR_GAP_CHAR",
    -5: "MH_ERR_ZERO_LENGTH",
    -6: "MH_ERR_UNSUPPORTED",
    -7: "MH_ERR_IO",
}

# Exported symbols of the C ABI (checked by the CPU test suite).
EXPORTS = [
    "mh_create", "mh_destroy", "mh_last_error", "mh_synchronize", "mh_kernel_time",
    "mh_kernel_time_reset", "mh_set_timing", "mh_get_seed", "mh_get_seed_length",
    "mh_get_seed_weight", "mh_default_seed_weight", "mh_sml_create", "mh_sml_create_device",
    "mh_sml_free", "mh_sml_get_header", "mh_sml_read", "mh_sml_get_seed_mers", "mh_sml_find_mer",
    "mh_sml_get_sequence", "mh_sml_save", "mh_sml_load", "mh_find_matches_pairwise", "mh_search_live_ranges",
    "mh_dp_align", "mh_dp_align_device", "mh_dp_align_device_noalloc", "mh_dp_noalloc_scratch_bytes",
    "mh_device_alloc", "mh_device_free", "mh_memcpy_h2d", "mh_memcpy_d2h",
    "mh_ascii_to_codes_device", "mh_chain_gap_jobs", "mh_find_matches_pairwise_list",
    "mh_find_matches_unique", "mh_match_k_list_size", "mh_match_k_list_genomes", "mh_match_k_list_starts",
    "mh_match_k_list_lengths", "mh_match_k_list_free",
    "mh_match_list_size", "mh_match_list_data", "mh_match_list_free", "mh_match_list_seq_count",
    "mh_match_list_seq_file", "mh_match_list_seq_length", "mh_match_list_write", "mh_match_list_read",
    "mh_index_from_xmfa", "mh_index_from_xmfa_file", "mh_index_fix_zero", "mh_index_fill_gaps", "mh_index_smooth_edges", "mh_xmfa_write",
    "mh_refine_pairwise", "mh_refine_anchor_cols", "mh_refine_pieces", "mh_anchor_gaps", "mh_chain_rows",
    "mh_recursion_gaps",
    "mh_dp_align_banded_device", "mh_dp_band_region", "mh_dp_band_region_ex", "mh_dp_band_cells",
    "mh_gap_features_device", "mh_mlp_create", "mh_mlp_free", "mh_mlp_set_params", "mh_mlp_get_params",
    "mh_mlp_forward", "mh_mlp_grad", "mh_mlp_train", "mh_band_predict",
]


class MhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class DpJob(C.Structure):
    _fields_ = [("a_off", C.c_uint64), ("b_off", C.c_uint64), ("a_len", C.c_uint32), ("b_len", C.c_uint32)]


class DpResult(C.Structure):
    _fields_ = [("score", C.c_int32), ("path_len", C.c_uint32), ("path_start", C.c_uint64)]


class BandStats(C.Structure):
    _fields_ = [("banded", C.c_uint64), ("rerun_doubled", C.c_uint64), ("rerun_full", C.c_uint64),
                ("cells", C.c_uint64)]


BAND_NO_FALLBACK = 1
MLP_IN, MLP_H1, MLP_H2 = 272, 256, 128
MLP_NPARAMS = MLP_H1 * MLP_IN + MLP_H1 + MLP_H2 * MLP_H1 + MLP_H2 + MLP_H2 + 1


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("weight_decay", C.c_float),
                ("seed", C.c_uint64)]


class Bmer(C.Structure):
    _fields_ = [("position", C.c_uint32), ("mer", C.c_uint64)]


class SmlHeader(C.Structure):
    _fields_ = [("version", C.c_uint32), ("alphabet_bits", C.c_uint32), ("seed", C.c_uint64),
                ("seed_length", C.c_uint32), ("seed_weight", C.c_uint32), ("length", C.c_uint64),
                ("sml_length", C.c_uint64), ("circular", C.c_uint8)]


class Match(C.Structure):
    _fields_ = [("start", C.c_int64 * MH_MAX_SEQ), ("length", C.c_uint64)]


class MatchStats(C.Structure):
    _fields_ = [("seed_tuples", C.c_int64), ("collisions", C.c_int64), ("matches", C.c_int64),
                ("repeat_runs", C.c_int64), ("walks", C.c_int64), ("long_walks", C.c_int64)]


DP_JOB_DTYPE = np.dtype([("a_off", "<u8"), ("b_off", "<u8"), ("a_len", "<u4"), ("b_len", "<u4")])
DP_RESULT_DTYPE = np.dtype([("score", "<i4"), ("path_len", "<u4"), ("path_start", "<u8")])
MATCH_DTYPE = np.dtype([("start", "<i8", (MH_MAX_SEQ,)), ("length", "<u8")])
BMER_DTYPE = np.dtype([("position", "<u4"), ("pad", "<u4"), ("mer", "<u8")])

_lib = None


def load_library(path=LIB_PATH):
    """Load libmauvehip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libmauvehip.so not built ({path}); run __graft_entry__.build()")
    L = C.CDLL(path)
    vp, u64, i64 = C.c_void_p, C.c_uint64, C.c_int64
    sig = {
        "mh_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "mh_destroy": (None, [vp]),
        "mh_last_error": (C.c_char_p, []),
        "mh_synchronize": (C.c_int, [vp]),
        "mh_kernel_time": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(i64)]),
        "mh_kernel_time_reset": (None, [vp]),
        "mh_set_timing": (C.c_int, [vp, C.c_int]),
        "mh_get_seed": (i64, [C.c_int, C.c_int]),
        "mh_get_seed_length": (C.c_int, [i64]),
        "mh_get_seed_weight": (C.c_int, [i64]),
        "mh_default_seed_weight": (C.c_uint32, [u64]),
        "mh_sml_create": (C.c_int, [vp, C.c_char_p, u64, u64, C.POINTER(vp)]),
        "mh_sml_create_device": (C.c_int, [vp, vp, u64, u64, C.POINTER(vp)]),
        "mh_sml_free": (None, [vp]),
        "mh_sml_get_header": (C.c_int, [vp, C.POINTER(SmlHeader)]),
        "mh_sml_read": (C.c_int, [vp, vp, u64, u64, vp, C.POINTER(u64)]),
        "mh_sml_get_seed_mers": (C.c_int, [vp, vp, vp, u64, vp]),
        "mh_sml_find_mer": (C.c_int, [vp, vp, u64, C.POINTER(u64), C.POINTER(C.c_int)]),
        "mh_sml_get_sequence": (C.c_int, [vp, vp, vp, u64]),
        "mh_sml_save": (C.c_int, [vp, vp, C.c_char_p]),
        "mh_sml_load": (C.c_int, [vp, C.c_char_p, C.POINTER(vp), C.POINTER(C.c_int)]),
        "mh_find_matches_pairwise": (C.c_int, [vp, C.POINTER(vp), C.c_int, vp, u64, C.POINTER(u64),
                                               C.POINTER(MatchStats)]),
        "mh_search_live_ranges": (C.c_int, [C.c_int, vp, vp, vp, u64, vp, u64, vp, vp]),
        "mh_dp_align": (C.c_int, [vp, vp, u64, vp, u64, vp, u64, vp, vp]),
        "mh_dp_align_device": (C.c_int, [vp, vp, vp, vp, u64, vp, vp]),
        "mh_dp_align_device_noalloc": (C.c_int, [vp, vp, vp, vp, u64, vp, vp]),
        "mh_dp_noalloc_scratch_bytes": (u64, []),
        "mh_dp_align_banded_device": (C.c_int, [vp, vp, vp, vp, u64, vp, C.c_uint32, vp, vp, vp,
                                                C.POINTER(BandStats)]),
        "mh_dp_band_region": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32)]),
        "mh_dp_band_region_ex": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "mh_dp_band_cells": (C.c_int, [vp, vp, vp, u64, vp]),
        "mh_gap_features_device": (C.c_int, [vp, vp, vp, vp, u64, vp]),
        "mh_mlp_create": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "mh_mlp_free": (None, [vp]),
        "mh_mlp_set_params": (C.c_int, [vp, vp, vp]),
        "mh_mlp_get_params": (C.c_int, [vp, vp, vp]),
        "mh_mlp_forward": (C.c_int, [vp, vp, vp, u64, vp]),
        "mh_mlp_grad": (C.c_int, [vp, vp, vp, vp, u64, vp, C.POINTER(C.c_float)]),
        "mh_mlp_train": (C.c_int, [vp, vp, vp, vp, u64, C.POINTER(MlpTrainOpts), vp]),
        "mh_band_predict": (C.c_int, [vp, vp, vp, u64, C.c_float, C.c_uint32, vp]),
        "mh_device_alloc": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "mh_device_free": (C.c_int, [vp, vp]),
        "mh_memcpy_h2d": (C.c_int, [vp, vp, vp, u64]),
        "mh_memcpy_d2h": (C.c_int, [vp, vp, vp, u64]),
        "mh_ascii_to_codes_device": (C.c_int, [vp, vp, u64, vp]),
        "mh_find_matches_pairwise_list": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.POINTER(vp),
                                                    C.POINTER(MatchStats)]),
        "mh_find_matches_unique": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.POINTER(vp), C.POINTER(MatchStats)]),
        "mh_match_k_list_size": (u64, [vp]),
        "mh_match_k_list_genomes": (C.c_int, [vp]),
        "mh_match_k_list_starts": (vp, [vp]),
        "mh_match_k_list_lengths": (vp, [vp]),
        "mh_match_k_list_free": (None, [vp]),
        "mh_match_list_size": (u64, [vp]),
        "mh_match_list_data": (vp, [vp]),
        "mh_match_list_free": (None, [vp]),
        "mh_match_list_seq_count": (C.c_int, [vp]),
        "mh_match_list_seq_file": (C.c_char_p, [vp, C.c_int]),
        "mh_match_list_seq_length": (u64, [vp, C.c_int]),
        "mh_match_list_write": (C.c_int, [C.c_char_p, vp, u64, C.c_int, C.POINTER(C.c_char_p), C.POINTER(u64)]),
        "mh_match_list_read": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "mh_index_from_xmfa": (C.c_int, [C.c_char_p, u64, u64, vp, C.POINTER(u64)]),
        "mh_index_from_xmfa_file": (C.c_int, [C.c_char_p, u64, vp, C.POINTER(u64)]),
        "mh_index_fix_zero": (C.c_int, [vp, u64, C.c_char_p, u64, C.c_char_p, u64]),
        "mh_index_fill_gaps": (C.c_int, [vp, u64, C.c_int]),
        "mh_index_smooth_edges": (C.c_int, [vp, u64, C.c_int]),
        "mh_xmfa_write": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.c_char_p, vp, u64]),
        "mh_refine_pairwise": (C.c_int, [vp, u64, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.POINTER(u64),
                                         C.c_uint32, C.POINTER(vp), C.POINTER(vp), C.POINTER(u64), C.POINTER(u64),
                                         C.POINTER(RefineStats)]),
        "mh_refine_anchor_cols": (C.c_int, [C.c_char_p, C.c_char_p, u64, vp, u64, C.POINTER(u64)]),
        "mh_refine_pieces": (C.c_int, [C.c_char_p, C.c_char_p, u64, vp, u64, C.POINTER(u64)]),
        "mh_anchor_gaps": (C.c_int, [vp, vp, u64, vp, u64, vp, u64, C.POINTER(vp), vp]),
        "mh_chain_rows": (C.c_int, [vp, vp, u64, C.c_char_p, u64, C.c_char_p, u64, vp, vp, u64, C.POINTER(u64)]),
        "mh_recursion_gaps": (C.c_int, [C.c_char_p, C.c_char_p, u64, i64, i64, u64, vp, u64, C.POINTER(u64)]),
        "mh_chain_gap_jobs": (C.c_int, [vp, u64, u64, u64, C.c_uint32, vp, C.POINTER(u64), vp, u64,
                                        C.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != MH_OK:
        msg = load_library().mh_last_error()
        raise MhError(rc, msg.decode() if msg else "")
    return rc


# ---- seeds (LM/SeedMasks.h) -------------------------------------------------

def get_seed(weight, rank=0):
    return load_library().mh_get_seed(weight, rank)


def get_seed_length(seed):
    return load_library().mh_get_seed_length(seed)


def get_seed_weight(seed):
    return load_library().mh_get_seed_weight(seed)


def get_default_seed_weight(avg_len):
    return load_library().mh_default_seed_weight(avg_len)


CODING_SEED = 3   # LM/SeedMasks.h:262


class Context:
    """One device and stream (mh_context)."""

    def __init__(self, device=0):
        L = load_library()
        h = C.c_void_p()
        _check(L.mh_create(device, C.byref(h)))
        self.h = h
        self.L = L

    def close(self):
        if self.h:
            self.L.mh_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(self.L.mh_synchronize(self.h))

    def set_timing(self, on=True):
        _check(self.L.mh_set_timing(self.h, 1 if on else 0))

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int64()
        _check(self.L.mh_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_kernel_times(self):
        self.L.mh_kernel_time_reset(self.h)

    # ---- device memory helpers ----
    def alloc(self, nbytes):
        p = C.c_void_p()
        _check(self.L.mh_device_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, ptr):
        _check(self.L.mh_device_free(self.h, ptr))

    def h2d(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        _check(self.L.mh_memcpy_h2d(self.h, dptr, arr.ctypes.data, arr.nbytes))

    def d2h(self, arr, dptr):
        _check(self.L.mh_memcpy_d2h(self.h, arr.ctypes.data, dptr, arr.nbytes))
        return arr


# ---- DP ----------------------------------------------------------------------

def make_jobs(a_off, a_len, b_off, b_len):
    jobs = np.zeros(len(a_len), dtype=DP_JOB_DTYPE)
    jobs["a_off"], jobs["a_len"] = a_off, a_len
    jobs["b_off"], jobs["b_len"] = b_off, b_len
    return jobs


def global_align(ctx, a_codes, b_codes, jobs):
    """Batch GlobalAlign.  Returns (paths buffer uint8, results structured array).

    a_codes / b_codes: uint8 letter codes 0..3; jobs: DP_JOB_DTYPE array.
    Job j's path is paths[r['path_start']: r['path_start'] + r['path_len']].
    """
    a = np.ascontiguousarray(a_codes, dtype=np.uint8)
    b = np.ascontiguousarray(b_codes, dtype=np.uint8)
    jobs = np.ascontiguousarray(jobs, dtype=DP_JOB_DTYPE)
    total = int((jobs["a_len"].astype(np.uint64) + jobs["b_len"]).sum())
    paths = np.zeros(max(total, 1), dtype=np.uint8)
    res = np.zeros(len(jobs), dtype=DP_RESULT_DTYPE)
    _check(ctx.L.mh_dp_align(ctx.h, a.ctypes.data, a.nbytes, b.ctypes.data, b.nbytes, jobs.ctypes.data,
                             len(jobs), paths.ctypes.data, res.ctypes.data))
    return paths, res


def path_of(paths, res, j):
    r = res[j]
    s = int(r["path_start"])
    return bytes(paths[s:s + int(r["path_len"])]).decode()


def global_align_device(ctx, d_a, d_b, jobs, d_paths, d_results):
    jobs = np.ascontiguousarray(jobs, dtype=DP_JOB_DTYPE)
    _check(ctx.L.mh_dp_align_device(ctx.h, d_a, d_b, jobs.ctypes.data, len(jobs), d_paths, d_results))


def global_align_banded_device(ctx, d_a, d_b, jobs, bands, d_paths, d_results, d_dev=None, flags=0):
    """Banded DP with edge fallback on device buffers; bands: uint32 per job
    (0 = full DP) or None.  Returns BandStats."""
    jobs = np.ascontiguousarray(jobs, dtype=DP_JOB_DTYPE)
    bp = None
    if bands is not None:
        bands = np.ascontiguousarray(bands, dtype=np.uint32)
        assert len(bands) == len(jobs)
        bp = bands.ctypes.data
    st = BandStats()
    _check(ctx.L.mh_dp_align_banded_device(ctx.h, d_a, d_b, jobs.ctypes.data, len(jobs), bp, flags, d_paths,
                                           d_results, d_dev, C.byref(st)))
    return st


def global_align_banded(ctx, a_codes, b_codes, jobs, bands, flags=0):
    """Host-buffer convenience over global_align_banded_device:
    (paths, results, dev, stats)."""
    a = np.ascontiguousarray(a_codes, dtype=np.uint8)
    b = np.ascontiguousarray(b_codes, dtype=np.uint8)
    jobs = np.ascontiguousarray(jobs, dtype=DP_JOB_DTYPE)
    n = len(jobs)
    total = int((jobs["a_len"].astype(np.uint64) + jobs["b_len"]).sum())
    bufs = [ctx.alloc(a.nbytes + 16), ctx.alloc(b.nbytes + 16), ctx.alloc(total + 16),
            ctx.alloc(16 * n + 16), ctx.alloc(4 * n + 16)]
    try:
        dA, dB, dP, dR, dD = bufs
        ctx.h2d(dA, a)
        ctx.h2d(dB, b)
        st = global_align_banded_device(ctx, dA, dB, jobs, bands, dP, dR, dD, flags)
        paths = np.zeros(max(total, 1), dtype=np.uint8)
        res = np.zeros(n, dtype=DP_RESULT_DTYPE)
        dev = np.zeros(n, dtype=np.uint32)
        ctx.d2h(paths, dP)
        ctx.d2h(res, dR)
        ctx.d2h(dev, dD)
    finally:
        for p in bufs:
            ctx.free(p)
    return paths, res, dev, st


REGION_FULL, REGION_STRIPES, REGION_WINDOW = 0, 1, 2


def band_cells(a_len, b_len, bands):
    """DP cells the scheduler evaluates per job at these bands (0 = full)."""
    la = np.ascontiguousarray(a_len, dtype=np.uint32)
    lb = np.ascontiguousarray(b_len, dtype=np.uint32)
    bd = np.ascontiguousarray(bands, dtype=np.uint32)
    out = np.zeros(len(la), dtype=np.uint64)
    _check(load_library().mh_dp_band_cells(la.ctypes.data, lb.ctypes.data, bd.ctypes.data, len(la),
                                           out.ctypes.data))
    return out


REGION_RERUN = 1


def band_region(a_len, b_len, band, rerun=False):
    """(effective band, region kind, width) the scheduler uses for a job; with
    rerun, the region of the edge fallback's second pass."""
    k, w = C.c_uint32(), C.c_uint32()
    eff = load_library().mh_dp_band_region_ex(int(a_len), int(b_len), int(band), REGION_RERUN if rerun else 0,
                                              C.byref(k), C.byref(w))
    return int(eff), int(k.value), int(w.value)


# ---- learned band ------------------------------------------------------------

def gap_features_device(ctx, d_a, d_b, jobs, d_feat):
    """float[len(jobs)][272] features into device buffer d_feat."""
    jobs = np.ascontiguousarray(jobs, dtype=DP_JOB_DTYPE)
    _check(ctx.L.mh_gap_features_device(ctx.h, d_a, d_b, jobs.ctypes.data, len(jobs), d_feat))


class Mlp:
    """The band network (mh_mlp): 272-256-128-1, bf16 weights, fp32 masters."""

    def __init__(self, ctx, seed=0, params=None):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.L.mh_mlp_create(ctx.h, seed, C.byref(h)))
        self.h = h.value
        if params is not None:
            self.set_params(params)

    def free(self):
        if self.h:
            self.ctx.L.mh_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def params(self):
        p = np.zeros(MLP_NPARAMS, dtype=np.float32)
        _check(self.ctx.L.mh_mlp_get_params(self.ctx.h, self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == MLP_NPARAMS
        _check(self.ctx.L.mh_mlp_set_params(self.ctx.h, self.h, p.ctypes.data))

    def forward_device(self, d_x, rows, d_y):
        _check(self.ctx.L.mh_mlp_forward(self.ctx.h, self.h, d_x, rows, d_y))

    def grad_device(self, d_x, d_t, rows):
        g = np.zeros(MLP_NPARAMS, dtype=np.float32)
        loss = C.c_float()
        _check(self.ctx.L.mh_mlp_grad(self.ctx.h, self.h, d_x, d_t, rows, g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def train_device(self, d_x, d_t, rows, epochs=4, batch=1024, lr=1e-3, weight_decay=0.0, seed=0):
        o = MlpTrainOpts(epochs, batch, lr, weight_decay, seed)
        curve = np.zeros(epochs, dtype=np.float32)
        _check(self.ctx.L.mh_mlp_train(self.ctx.h, self.h, d_x, d_t, rows, C.byref(o), curve.ctypes.data))
        return curve

    def predict_bands_device(self, d_feat, rows, delta, min_band=4):
        b = np.zeros(rows, dtype=np.uint32)
        _check(self.ctx.L.mh_band_predict(self.ctx.h, self.h, d_feat, rows, float(delta), int(min_band),
                                          b.ctypes.data))
        return b


# ---- sorted mer lists (DNAFileSML / DNAMemorySML) -------------------------

class SortedMerList:
    """A device-resident sorted mer list (mh_sml).

    ``SortedMerList(ctx, seq, seed)`` builds the list as FileSML::Create does
    (LM/FileSML.cpp:401-459); ``seq`` is ASCII bytes / str / uint8 array.
    """

    def __init__(self, ctx, seq, seed):
        if isinstance(seq, str):
            seq = seq.encode()
        if isinstance(seq, np.ndarray):
            seq = np.ascontiguousarray(seq, dtype=np.uint8).tobytes()
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.L.mh_sml_create(ctx.h, seq, len(seq), seed, C.byref(h)))
        self.h = h

    @classmethod
    def from_device(cls, ctx, d_seq, length, seed):
        self = cls.__new__(cls)
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.L.mh_sml_create_device(ctx.h, d_seq, length, seed, C.byref(h)))
        self.h = h
        return self

    def free(self):
        if getattr(self, "h", None):
            self.ctx.L.mh_sml_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def header(self):
        hd = SmlHeader()
        _check(self.ctx.L.mh_sml_get_header(self.h, C.byref(hd)))
        return hd

    def sml_length(self):
        return self.header().sml_length

    def read(self, offset=0, count=None):
        """Entries [offset, offset+count): (positions uint32, mers uint64)."""
        n = self.sml_length()
        if count is None:
            count = max(n - offset, 0)
        buf = np.zeros(max(count, 1), dtype=BMER_DTYPE)
        got = C.c_uint64()
        _check(self.ctx.L.mh_sml_read(self.ctx.h, self.h, offset, count, buf.ctypes.data, C.byref(got)))
        buf = buf[:got.value]
        return buf["position"].copy(), buf["mer"].copy()

    def get_seed_mers(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.uint32)
        out = np.zeros(len(pos), dtype=np.uint64)
        _check(self.ctx.L.mh_sml_get_seed_mers(self.ctx.h, self.h, pos.ctypes.data, len(pos), out.ctypes.data))
        return out

    def find_mer(self, mer):
        idx, found = C.c_uint64(), C.c_int()
        _check(self.ctx.L.mh_sml_find_mer(self.ctx.h, self.h, mer, C.byref(idx), C.byref(found)))
        return bool(found.value), idx.value

    def sequence_words(self):
        hd = self.header()
        nw = (hd.length * 2 + 31) // 32 + 2
        out = np.zeros(nw, dtype=np.uint32)
        _check(self.ctx.L.mh_sml_get_sequence(self.ctx.h, self.h, out.ctypes.data, nw))
        return out

    def save(self, path):
        """Write the list as an .sslist file (the file FileSML::Create leaves,
        LM/FileSML.cpp:427-446)."""
        _check(self.ctx.L.mh_sml_save(self.ctx.h, self.h, os.fsencode(path)))

    @classmethod
    def load(cls, ctx, path):
        """Read an .sslist file (FileSML::LoadFile, LM/FileSML.cpp:48-118).
        Raises SmlLoadError carrying LoadFile2's status code (1 cannot open,
        2 short header, 3 other format version, 4 short sequence, 5 short
        position list; 6 header this build does not represent, 7 not a sorted
        list)."""
        h = C.c_void_p()
        st = C.c_int()
        rc = ctx.L.mh_sml_load(ctx.h, os.fsencode(path), C.byref(h), C.byref(st))
        if rc != MH_OK:
            msg = ctx.L.mh_last_error()
            raise SmlLoadError(rc, st.value, msg.decode() if msg else "")
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = h
        return self

    def seed(self):
        return self.header().seed


class SmlLoadError(MhError):
    def __init__(self, code, status, msg):
        super().__init__(code, f"status {status}: {msg}")
        self.status = status


def pattern_text(seed):
    """The seed pattern as text, '1'/'0' from its highest set bit down
    (getPatternText, MA/progressiveMauve.cpp:198-213)."""
    return format(int(seed) & 0xFFFFFFFFFFFFFFFF, "b")


def default_sml_file_names(seq_files, seed_weight, seed_rank=0):
    """<sequence file>.<seed pattern>.sslist for each sequence file
    (getDefaultSmlFileNames, MA/progressiveMauve.cpp:215-225)."""
    pat = pattern_text(get_seed(seed_weight, seed_rank))
    return [f"{f}.{pat}.sslist" for f in seq_files]


def load_smls(ctx, seqs, sml_files, seed, force_create=False, log=None):
    """GenericMatchList::LoadSMLs (LM/MatchList.h:264-378) for sequences
    given as ASCII: each list is loaded from its file; a list whose file is
    missing or unreadable, whose seed differs from ``seed``, or every list
    when ``force_create``, is built on the device and written to its file.
    Returns the lists (device-resident) in sequence order."""
    out = []
    for seq, path in zip(seqs, sml_files):
        sml = None
        if not force_create:
            try:
                sml = SortedMerList.load(ctx, path)
            except SmlLoadError:
                sml = None
            if sml is not None and sml.seed() != seed:
                if log:
                    log("Default seed mismatch.  A new sorted mer list will be created.")
                sml.free()
                sml = None
            elif sml is not None and log:
                log("Sorted mer list loaded successfully")
        if sml is None:
            if log:
                log("Creating sorted mer list")
            sml = SortedMerList(ctx, seq, seed)
            sml.save(path)
        out.append(sml)
    return out


def find_matches_raw(ctx, smls):
    """PairwiseMatchFinder::FindMatches as one mh_match array (MATCH_DTYPE) and MatchStats."""
    k = len(smls)
    arr = (C.c_void_p * k)(*[s.h.value for s in smls])
    st = MatchStats()
    h = C.c_void_p()
    _check(ctx.L.mh_find_matches_pairwise_list(ctx.h, arr, k, C.byref(h), C.byref(st)))
    try:
        n = ctx.L.mh_match_list_size(h)
        out = np.empty(n, dtype=MATCH_DTYPE)
        if n:
            C.memmove(out.ctypes.data, ctx.L.mh_match_list_data(h), n * MATCH_DTYPE.itemsize)
    finally:
        ctx.L.mh_match_list_free(h)
    return out, st


def find_matches_unique(ctx, smls):
    """UniqueMatchFinder::FindMatches over 2 to 16 sorted mer lists:
    (starts (n, k) int64, lengths uint64, MatchStats)."""
    k = len(smls)
    arr = (C.c_void_p * k)(*[s.h.value for s in smls])
    st = MatchStats()
    h = C.c_void_p()
    _check(ctx.L.mh_find_matches_unique(ctx.h, arr, k, C.byref(h), C.byref(st)))
    try:
        n = ctx.L.mh_match_k_list_size(h)
        starts = np.zeros((n, k), dtype=np.int64)
        lengths = np.zeros(n, dtype=np.uint64)
        if n:
            C.memmove(starts.ctypes.data, ctx.L.mh_match_k_list_starts(h), n * k * 8)
            C.memmove(lengths.ctypes.data, ctx.L.mh_match_k_list_lengths(h), n * 8)
    finally:
        ctx.L.mh_match_k_list_free(h)
    return starts, lengths, st


def search_live_ranges(keys, lens, seed):
    """Host-only plan of SearchRange's repeat-limit restarts over sorted
    64-bit key arrays (one per list) of sequences of the given lengths:
    (list of [(lo, hi), ...] per list, number of restarts)."""
    L = load_library()
    k = len(keys)
    ks = [np.ascontiguousarray(x, dtype=np.uint64) for x in keys]
    ptrs = (C.c_void_p * k)(*[x.ctypes.data if len(x) else None for x in ks])
    n = (C.c_uint64 * k)(*[len(x) for x in ks])
    ln = (C.c_uint64 * k)(*[int(x) for x in lens])
    cap = 64
    while True:
        buf = np.zeros(3 * cap, dtype=np.uint64)
        nr = C.c_uint64()
        rs = C.c_int64()
        rc = L.mh_search_live_ranges(k, ptrs, n, ln, C.c_uint64(seed), buf.ctypes.data, cap, C.byref(nr),
                                     C.byref(rs))
        if rc == -1 and nr.value > cap:
            cap = nr.value
            continue
        _check(rc)
        break
    out = [[] for _ in range(k)]
    for t in range(nr.value):
        out[int(buf[3 * t])].append((int(buf[3 * t + 1]), int(buf[3 * t + 2])))
    return out, rs.value


def find_matches(ctx, smls):
    """PairwiseMatchFinder::FindMatches: (starts int64 (n, k), lengths uint64, MatchStats)."""
    out, st = find_matches_raw(ctx, smls)
    return out["start"][:, :len(smls)].copy(), out["length"].copy(), st


ANCHOR_GAP_DTYPE = np.dtype([("lend", "<i8", (2,)), ("rend", "<i8", (2,))])


def anchor_gaps_device(ctx, d_a, len_a, d_b, len_b, gaps):
    """pairwiseAnchorSearch over gaps (ANCHOR_GAP_DTYPE records) of two device
    ASCII genomes: (starts (n, 2), lengths, first (n_gaps + 1))."""
    g = np.ascontiguousarray(gaps, dtype=ANCHOR_GAP_DTYPE)
    first = np.zeros(len(g) + 1, dtype=np.uint64)
    h = C.c_void_p()
    _check(ctx.L.mh_anchor_gaps(ctx.h, d_a, len_a, d_b, len_b, g.ctypes.data if len(g) else None, len(g),
                                C.byref(h), first.ctypes.data))
    try:
        n = ctx.L.mh_match_list_size(h)
        out = np.empty(n, dtype=MATCH_DTYPE)
        if n:
            C.memmove(out.ctypes.data, ctx.L.mh_match_list_data(h), n * MATCH_DTYPE.itemsize)
    finally:
        ctx.L.mh_match_list_free(h)
    return out["start"][:, :2].copy(), out["length"].copy(), first


def chain_rows(matches, chain, seq_a, seq_b, as_array=False):
    """mh_chain_rows: (row_a, row_b) as bytes (uint8 arrays if as_array) for a
    chain of MATCH_DTYPE matches (chain = indices, as chain_gap_jobs_raw
    returns them)."""
    L = load_library()
    m = np.ascontiguousarray(matches, dtype=MATCH_DTYPE)
    ch = np.ascontiguousarray(chain, dtype=np.uint32)
    a = seq_a if isinstance(seq_a, bytes) else np.ascontiguousarray(seq_a, dtype=np.uint8).tobytes()
    b = seq_b if isinstance(seq_b, bytes) else np.ascontiguousarray(seq_b, dtype=np.uint8).tobytes()
    cols = C.c_uint64()
    cap = len(a) + len(b)
    ra = np.empty(cap + 1, dtype=np.uint8)
    rb = np.empty(cap + 1, dtype=np.uint8)
    _check(L.mh_chain_rows(m.ctypes.data, ch.ctypes.data, len(ch), a, len(a), b, len(b), ra.ctypes.data,
                           rb.ctypes.data, cap, C.byref(cols)))
    n = cols.value
    if as_array:
        return ra[:n], rb[:n]
    return ra[:n].tobytes(), rb[:n].tobytes()


def recursion_gaps(row_a, row_b, left_a=1, left_b=1, min_gap=200):
    """mh_recursion_gaps: ANCHOR_GAP_DTYPE records of the gaps searched recursively."""
    L = load_library()
    (a, na, ka), (b, nb, kb) = _row_ptr(row_a), _row_ptr(row_b)
    if na != nb:
        raise ValueError("rows differ in length")
    n = C.c_uint64()
    # each gap spans more than min_gap columns: this many always suffice
    out = np.zeros(na // (min_gap + 1) + 2, dtype=ANCHOR_GAP_DTYPE)
    _check(L.mh_recursion_gaps(a, b, na, left_a, left_b, min_gap, out.ctypes.data, len(out), C.byref(n)))
    return out[:n.value]


def anchor_gaps(ctx, seq_a, seq_b, gaps):
    """anchor_gaps_device with host ASCII genomes."""
    a = seq_a.encode() if isinstance(seq_a, str) else bytes(seq_a)
    b = seq_b.encode() if isinstance(seq_b, str) else bytes(seq_b)
    da, db = ctx.alloc(len(a) + 16), ctx.alloc(len(b) + 16)
    try:
        ctx.h2d(da, np.frombuffer(a, dtype=np.uint8))
        ctx.h2d(db, np.frombuffer(b, dtype=np.uint8))
        return anchor_gaps_device(ctx, da, len(a), db, len(b), gaps)
    finally:
        ctx.free(da)
        ctx.free(db)


class RefineStats(C.Structure):
    _fields_ = [("profile_cells", C.c_uint64), ("window_cells", C.c_uint64)]


REFINE_PROFILE = 1
REFINE_WINDOWS = 2


def _row_bytes(r):
    return r.encode() if isinstance(r, str) else bytes(r)


def _row_ptr(r):
    """(char pointer, length, object to keep alive) for a row given as str,
    bytes or a contiguous uint8 array (arrays are passed without a copy)."""
    if isinstance(r, np.ndarray) and r.dtype == np.uint8 and r.ndim == 1 and r.flags.c_contiguous:
        return C.c_char_p(r.ctypes.data if len(r) else None), len(r), r
    b = _row_bytes(r)
    return b, len(b), b


def refine_anchor_cols(row1, row2):
    """FindAnchorColsPP over two aligned rows (host part of mh_refine_pairwise)."""
    L = load_library()
    a, b = _row_bytes(row1), _row_bytes(row2)
    if len(a) != len(b):
        raise ValueError("rows differ in length")
    out = np.zeros(max(1, len(a)), dtype=np.uint32)
    n = C.c_uint64()
    _check(L.mh_refine_anchor_cols(a, b, len(a), out.ctypes.data, len(out), C.byref(n)))
    return out[:n.value].tolist()


def refine_pieces(row1, row2):
    """The pieces refineAlignment cuts two rows into: [(lo, hi, unaligned)]."""
    L = load_library()
    a, b = _row_bytes(row1), _row_bytes(row2)
    if len(a) != len(b):
        raise ValueError("rows differ in length")
    cap = len(a) + 2
    out = np.zeros(3 * cap, dtype=np.uint64)
    n = C.c_uint64()
    _check(L.mh_refine_pieces(a, b, len(a), out.ctypes.data, cap, C.byref(n)))
    return [(int(out[3 * k]), int(out[3 * k + 1]), bool(out[3 * k + 2])) for k in range(n.value)]


def refine_pairwise(ctx, intervals, passes=REFINE_PROFILE | REFINE_WINDOWS, stats=None, as_array=False):
    """mh_refine_pairwise over [(row1, row2)]: the refined row pairs (bytes;
    uint8 arrays if as_array).  A RefineStats passed as stats receives the DP
    cells of each pass."""
    L = load_library()
    n = len(intervals)
    r1 = [_row_ptr(a) for a, _ in intervals]
    r2 = [_row_ptr(b) for _, b in intervals]
    for a, b in zip(r1, r2):
        if a[1] != b[1]:
            raise ValueError("rows of an interval differ in length")
    cols = (C.c_uint64 * max(1, n))(*[a[1] for a in r1])
    caps = (C.c_uint64 * max(1, n))(*[2 * a[1] + 1 for a in r1])
    o1 = [np.empty(2 * a[1] + 1, dtype=np.uint8) for a in r1]
    o2 = [np.empty(2 * a[1] + 1, dtype=np.uint8) for a in r1]
    outc = (C.c_uint64 * max(1, n))()
    _check(L.mh_refine_pairwise(ctx.h, n, (C.c_char_p * max(1, n))(*[a[0] for a in r1]),
                                (C.c_char_p * max(1, n))(*[b[0] for b in r2]), cols,
                                passes, (C.c_void_p * max(1, n))(*[b.ctypes.data for b in o1]),
                                (C.c_void_p * max(1, n))(*[b.ctypes.data for b in o2]), caps, outc,
                                C.byref(stats) if stats is not None else None))
    if as_array:
        return [(o1[i][:outc[i]], o2[i][:outc[i]]) for i in range(n)]
    return [(o1[i][:outc[i]].tobytes(), o2[i][:outc[i]].tobytes()) for i in range(n)]


class XmfaRow(C.Structure):
    _fields_ = [("start", C.c_int64), ("length", C.c_uint64), ("row", C.c_char_p), ("row_len", C.c_uint64)]


def write_xmfa(path, intervals, seq_files, backbone_file=None):
    """WriteStandardAlignment (LM/IntervalList.h:353-442).  ``intervals`` is a
    list of intervals, each a list (one per genome) of (start, length, row):
    start 1-based, negative on the reverse strand, 0 if absent; row the
    aligned string."""
    L = load_library()
    k = len(seq_files)
    rows = (XmfaRow * max(1, k * len(intervals)))()
    keep = []
    for i, iv in enumerate(intervals):
        if len(iv) != k:
            raise ValueError("every interval needs one row per genome")
        for s, (start, length, row) in enumerate(iv):
            b, nb, kb = _row_ptr(row)
            keep.append(kb)
            rows[i * k + s] = XmfaRow(int(start), int(length), b, nb)
    files = (C.c_char_p * k)(*[os.fsencode(f) for f in seq_files])
    _check(L.mh_xmfa_write(os.fsencode(path), k, files, os.fsencode(backbone_file) if backbone_file else None,
                           C.cast(rows, C.c_void_p), len(intervals)))


def write_match_list(path, starts, lengths, seq_files=None, seq_lengths=None):
    """WriteList (LM/MatchList.h:617-662): the `--mums` match-list file for
    matches given as starts (n, k) int64 and lengths (n,)."""
    L = load_library()
    starts = np.asarray(starts, dtype=np.int64)
    k = starts.shape[1] if starts.ndim == 2 else len(seq_files or ())
    n = len(lengths)
    m = np.zeros(n, dtype=MATCH_DTYPE)
    if n:
        m["start"][:, :k] = starts
        m["length"] = np.asarray(lengths, dtype=np.uint64)
    files = None
    if seq_files is not None:
        files = (C.c_char_p * k)(*[None if f is None else os.fsencode(f) for f in seq_files])
    lens = None
    if seq_lengths is not None:
        lens = (C.c_uint64 * k)(*[int(v) for v in seq_lengths])
    _check(L.mh_match_list_write(os.fsencode(path), m.ctypes.data, n, k, files, lens))


def read_match_list(path):
    """ReadList (LM/MatchList.h:524-613): (starts (n, k) int64, lengths
    uint64, sequence file names, sequence lengths)."""
    L = load_library()
    h = C.c_void_p()
    _check(L.mh_match_list_read(os.fsencode(path), C.byref(h)))
    try:
        k = L.mh_match_list_seq_count(h)
        n = L.mh_match_list_size(h)
        out = np.empty(n, dtype=MATCH_DTYPE)
        if n:
            C.memmove(out.ctypes.data, L.mh_match_list_data(h), n * MATCH_DTYPE.itemsize)
        files = [L.mh_match_list_seq_file(h, i).decode() for i in range(k)]
        lens = [L.mh_match_list_seq_length(h, i) for i in range(k)]
    finally:
        L.mh_match_list_free(h)
    return out["start"][:, :k].copy(), out["length"].copy(), files, lens


def ascii_to_codes_device(ctx, d_ascii, length, d_codes):
    _check(ctx.L.mh_ascii_to_codes_device(ctx.h, d_ascii, length, d_codes))


def chain_gap_jobs(starts, lengths, len_a, len_b, max_gap=20000):
    """Stand-in DP job producer (see mh_chain_gap_jobs): (chain indices, jobs)."""
    n = len(lengths)
    m = np.zeros(n, dtype=MATCH_DTYPE)
    if n:
        m["start"][:, :starts.shape[1]] = starts
        m["length"] = lengths
    return chain_gap_jobs_raw(m, len_a, len_b, max_gap)


def chain_gap_jobs_raw(matches, len_a, len_b, max_gap=20000):
    """chain_gap_jobs on an mh_match array (as find_matches_raw returns it)."""
    L = load_library()
    m = np.ascontiguousarray(matches, dtype=MATCH_DTYPE)
    n = len(m)
    chain = np.empty(max(n, 1), dtype=np.uint32)
    nc, nj = C.c_uint64(), C.c_uint64()
    jobs = np.empty(n + 1, dtype=DP_JOB_DTYPE)    # a chain of k anchors has at most k + 1 gaps
    _check(L.mh_chain_gap_jobs(m.ctypes.data if n else None, n, len_a, len_b, max_gap, chain.ctypes.data,
                               C.byref(nc), jobs.ctypes.data, n + 1, C.byref(nj)))
    return chain[:nc.value], jobs[:nj.value]
