"""ctypes mirror of include/gd_remc.h and thin Python helpers.

The product is the C ABI in libgdremc.so; this module only marshals numpy
arrays into the ABI structs for tests, the benchmark and Python callers.
Every compute call goes to the HIP library; when the library is missing the
import of `lib()` raises instead of falling back to anything.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgdremc.so")

MAXTP1, MAXTP2, MAXTP3, MAXTP4, MAXWEI = 30, 24, 50, 20, 10
MAX_LNA = 128
MCS_INVALID = 728492.0
TERM_ORDER_FILE, TERM_ORDER_SLOT = 0, 1

# Energy slots (reference Energy.e[] order, dock.h:100-115)
SLOTS = ("vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total")
# Parameter-file order (weights, nor_a, nor_b): vdw ele pmf hpc hdb dst psp kde lhm
FILE_ORDER = ("vdw", "ele", "pmf", "hpc", "hdb", "dst", "psp", "kde", "lhm")
FILE_TO_SLOT = [SLOTS.index(n) for n in FILE_ORDER]

F = C.c_float
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int)


class ComplexSize(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("n_lig", "n_prt", "n_tmp", "n_rep", "lna", "pnp", "pnk", "pos")]


class Ligand(C.Structure):
    _fields_ = [("x", FP), ("y", FP), ("z", FP), ("t", IP), ("c", FP), ("n", IP),
                ("center", F * 3)]


class Protein(C.Structure):
    _fields_ = [("x", FP), ("y", FP), ("z", FP), ("t", IP), ("c", IP), ("ele", FP),
                ("seq3r", IP), ("c0_and_d12_or_c2", IP), ("hpp", FP),
                ("pocket_center", F * 3)]


class Psp(C.Structure):
    _fields_ = [("psp", FP), ("n_res", C.c_int)]


class Kde(C.Structure):
    _fields_ = [("x", FP), ("y", FP), ("z", FP), ("t", IP)]


class Mcs(C.Structure):
    _fields_ = [("x", FP), ("y", FP), ("z", FP), ("tcc", FP), ("n_field", C.c_int)]


T2x1 = F * MAXTP1 * MAXTP2


class EnePara(C.Structure):
    _fields_ = [("p1a", T2x1), ("p2a", T2x1), ("lj0", F), ("lj1", F), ("el0", F),
                ("el1", F), ("ele", F * MAXTP3), ("a1", F), ("b1", F),
                ("pmf0", T2x1), ("pmf1", T2x1), ("hdb0", T2x1), ("hdb1", T2x1),
                ("hdb2", T2x1), ("hpp", F * MAXTP4), ("hpl0", F * MAXTP2),
                ("hpl1", F * MAXTP2), ("hpl2", F * MAXTP2), ("kde2", F), ("kde3", F),
                ("w", F * MAXWEI), ("a_para", F * MAXWEI), ("b_para", F * MAXWEI)]


T1x2x2 = F * 2 * MAXTP2 * MAXTP1


class EnePara0(C.Structure):
    _fields_ = [("vdw", T1x2x2), ("ele", F * MAXTP3), ("pmf", T1x2x2),
                ("hpp", F * MAXTP4), ("hpl", F * 2 * MAXTP2), ("hdb", T1x2x2),
                ("lj", F * 3), ("el", F * 2), ("kde", F), ("w", F * MAXWEI),
                ("a_para", F * MAXWEI), ("b_para", F * MAXWEI)]


class Temp(C.Structure):
    _fields_ = [("t", F), ("minus_beta", F), ("order", C.c_int)]


class Replica(C.Structure):
    _fields_ = [("idx_rep", C.c_int), ("idx_prt", C.c_int), ("idx_tmp", C.c_int),
                ("idx_lig", C.c_int)]


class McPara(C.Structure):
    _fields_ = [("steps_total", C.c_int), ("steps_per_dump", C.c_int),
                ("steps_per_exchange", C.c_int), ("move_scale", F * 6)]


class McLog(C.Structure):
    _fields_ = [("t0", C.c_double), ("t1", C.c_double), ("t2", C.c_double),
                ("ac_mc", C.c_longlong), ("acs_mc", IP), ("ac_temp_exchg", C.c_longlong),
                ("acs_temp_exchg", IP)]


class Options(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("complex_id", C.c_uint32), ("exchange", C.c_int),
                ("away", C.c_int), ("term_order", C.c_int), ("device", C.c_int),
                ("record_stride", C.c_int), ("record_prefix", C.c_char_p),
                ("mcc", C.c_int), ("combine", C.c_int), ("fixed_draws", C.c_int),
                ("reverse", C.c_int)]


class StepRecord(C.Structure):
    """gd_step_record (LigRecordSingleStep, dock.h:487-494)."""
    _fields_ = [("replica", Replica), ("energy", F * (MAXWEI + 1)), ("movematrix", F * 6),
                ("step", C.c_int)]


STEP_RECORD_DTYPE = np.dtype([("idx_rep", "<i4"), ("idx_prt", "<i4"), ("idx_tmp", "<i4"),
                              ("idx_lig", "<i4"), ("energy", "<f4", (MAXWEI + 1,)),
                              ("movematrix", "<f4", (6,)), ("step", "<i4")])


class ReplicaResult(C.Structure):
    _fields_ = [("move_final", F * 6), ("energy_final", F * MAXWEI),
                ("move_best", F * 6), ("energy_best", F * MAXWEI),
                ("energy_best_fp64", C.c_double * MAXWEI), ("step_best", C.c_int),
                ("idx_tmp_final", C.c_int), ("acs_mc", C.c_int), ("acs_exchg", C.c_int),
                ("cmcc_final", F), ("cmcc_best", F)]


class Job(C.Structure):
    _fields_ = [("lig", C.POINTER(Ligand)), ("prt", C.POINTER(Protein)),
                ("psp", C.POINTER(Psp)), ("kde", C.POINTER(Kde)), ("mcs", C.POINTER(Mcs)),
                ("enepara", C.POINTER(EnePara)), ("temp", C.POINTER(Temp)),
                ("replica", C.POINTER(Replica)), ("mcpara", C.POINTER(McPara)),
                ("size", ComplexSize), ("opt", C.POINTER(Options))]


class EngineConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("arena_bytes", C.c_size_t), ("wpg", C.c_int),
                ("persistent", C.c_int), ("use_graph", C.c_int), ("kde_policy", C.c_int),
                ("staging_bytes", C.c_size_t), ("occ", C.c_int), ("split", C.c_int),
                ("prio", C.c_int)]


TUNE_MAX = 32


class TuneReport(C.Structure):
    _fields_ = [("n", C.c_int), ("wpg", C.c_int * TUNE_MAX),
                ("persistent", C.c_int * TUNE_MAX), ("kde_global", C.c_int * TUNE_MAX),
                ("occ", C.c_int * TUNE_MAX), ("ms", C.c_double * TUNE_MAX),
                ("split", C.c_int * TUNE_MAX),
                ("chosen_wpg", C.c_int), ("chosen_persistent", C.c_int),
                ("chosen_kde_global", C.c_int), ("chosen_occ", C.c_int),
                ("chosen_split", C.c_int)]


class EngineLayout(C.Structure):
    _fields_ = [("stage_bytes", C.c_int), ("wpg", C.c_int), ("persistent", C.c_int),
                ("n_tasks", C.c_int), ("n_chains", C.c_int), ("n_complexes", C.c_int),
                ("grid", C.c_int), ("resident_groups", C.c_int), ("kde_in_lds", C.c_int),
                ("arena_used", C.c_size_t), ("arena_size", C.c_size_t),
                ("n_classes", C.c_int), ("kde_global", C.c_int), ("occ", C.c_int),
                ("split", C.c_int)]


class BatchConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("wpg", C.c_int), ("persistent", C.c_int),
                ("no_graph", C.c_int), ("autotune", C.c_int), ("kde_policy", C.c_int),
                ("occ", C.c_int), ("split", C.c_int)]


class BatchReport(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("tune_s", C.c_double), ("run_s", C.c_double),
                ("results_s", C.c_double), ("cycles", C.c_int), ("tune", TuneReport),
                ("layout", EngineLayout)]


SERVER_MAX_LANES = 4   # GD_SERVER_MAX_LANES


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("arena_bytes", C.c_size_t), ("wpg", C.c_int),
                ("persistent", C.c_int), ("use_graph", C.c_int), ("kde_policy", C.c_int),
                ("max_resident_replicas", C.c_int), ("cycles_per_tick", C.c_int),
                ("max_ticks_in_flight", C.c_int), ("keep_results", C.c_int),
                ("readback_bytes", C.c_size_t), ("occ", C.c_int), ("split", C.c_int),
                ("lanes", C.c_int), ("lane_max_pairs", C.c_int * (SERVER_MAX_LANES - 1)),
                ("lane_split", C.c_int * SERVER_MAX_LANES),
                ("lane_max_jobs", C.c_int * SERVER_MAX_LANES),
                ("lane_prio", C.c_int * SERVER_MAX_LANES),
                ("lane_max_total", C.c_int * SERVER_MAX_LANES)]


class ServerJobInfo(C.Structure):
    _fields_ = [("ticket", C.c_longlong), ("state", C.c_int), ("status", C.c_int),
                ("n_rep", C.c_int), ("steps_total", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("latency", C.c_double),
                ("footprint", C.c_size_t), ("lane", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong),
                ("failed", C.c_longlong), ("cancelled", C.c_longlong), ("queued", C.c_int),
                ("running", C.c_int), ("p50_latency", C.c_double),
                ("p99_latency", C.c_double), ("mean_latency", C.c_double),
                ("max_latency", C.c_double), ("p50_wait", C.c_double),
                ("p99_wait", C.c_double), ("pose_evals", C.c_double),
                ("first_submit", C.c_double), ("last_done", C.c_double),
                ("pose_evals_per_s", C.c_double), ("arena_size", C.c_size_t),
                ("arena_peak", C.c_size_t), ("running_peak", C.c_int),
                ("cycles", C.c_longlong), ("rebuilds", C.c_longlong), ("lanes", C.c_int),
                ("lane_cycles", C.c_longlong * SERVER_MAX_LANES),
                ("lane_completed", C.c_longlong * SERVER_MAX_LANES),
                ("lane_split", C.c_int * SERVER_MAX_LANES)]


class InputFiles(C.Structure):
    _fields_ = [("ligand_sdf", C.c_char_p), ("molid", C.c_char_p),
                ("protein_pdb", C.c_char_p), ("lhm_ff", C.c_char_p),
                ("ligand_id", C.c_char_p), ("flags", C.c_int)]


LOAD_PLATE_ALL_MODELS = 1


class InputsView(C.Structure):
    _fields_ = [("size", ComplexSize), ("lig", Ligand), ("prt", Protein), ("psp", Psp),
                ("kde", Kde), ("mcs", Mcs), ("lens_rmsd", FP), ("prt_d", IP),
                ("prt_r", IP), ("ligand_name", C.c_char_p)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)

_LIB = None

SYMBOLS = {
    "gd_run": (C.c_int, [C.POINTER(Ligand), C.POINTER(Protein), C.POINTER(Psp),
                         C.POINTER(Kde), C.POINTER(Mcs), C.POINTER(EnePara),
                         C.POINTER(Temp), C.POINTER(Replica), C.POINTER(McPara),
                         C.POINTER(McLog), ComplexSize, C.POINTER(Options),
                         C.POINTER(ReplicaResult)]),
    "gd_score_poses": (C.c_int, [C.POINTER(Ligand), C.POINTER(Protein), C.POINTER(Psp),
                                 C.POINTER(Kde), C.POINTER(Mcs), C.POINTER(EnePara),
                                 ComplexSize, C.c_int, IP, IP, FP, C.c_int, C.c_int, FP,
                                 C.POINTER(C.c_double)]),
    "gd_score_poses_noalloc": (C.c_int, [C.POINTER(Ligand), C.POINTER(Protein),
                                         C.POINTER(Psp), C.POINTER(Kde), C.POINTER(Mcs),
                                         C.POINTER(EnePara), ComplexSize, C.c_int, IP, IP, FP,
                                         C.c_int, C.c_int, FP, C.POINTER(C.c_double)]),
    "gd_session_create": (C.c_int, [C.POINTER(Ligand), C.POINTER(Protein),
                                    C.POINTER(Psp), C.POINTER(Kde), C.POINTER(Mcs),
                                    C.POINTER(EnePara), C.POINTER(Temp),
                                    C.POINTER(Replica), C.POINTER(McPara), ComplexSize,
                                    C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "gd_session_init": (C.c_int, [C.c_void_p]),
    "gd_session_cycles": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_session_sync": (C.c_int, [C.c_void_p]),
    "gd_session_steps_done": (C.c_longlong, [C.c_void_p]),
    "gd_session_results": (C.c_int, [C.c_void_p, C.POINTER(McLog),
                                     C.POINTER(ReplicaResult)]),
    "gd_session_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_session_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]),
    "gd_session_layout": (C.c_int, [C.c_void_p, IP, IP, IP, IP]),
    "gd_session_records": (C.c_int, [C.c_void_p, IP, IP, C.c_void_p]),
    "gd_session_destroy": (None, [C.c_void_p]),
    "gd_engine_create": (C.c_int, [C.POINTER(EngineConfig), C.POINTER(C.c_void_p)]),
    "gd_engine_destroy": (None, [C.c_void_p]),
    "gd_job_footprint": (C.c_int, [C.POINTER(Job), C.c_int, C.POINTER(C.c_size_t)]),
    "gd_engine_add": (C.c_int, [C.c_void_p, C.POINTER(Job), IP]),
    "gd_engine_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_engine_cycles": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_engine_sync": (C.c_int, [C.c_void_p]),
    "gd_engine_cycles_left": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_engine_results": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(McLog),
                                    C.POINTER(ReplicaResult)]),
    "gd_engine_autotune": (C.c_int, [C.c_void_p, C.POINTER(TuneReport)]),
    "gd_engine_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_engine_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "gd_engine_layout": (C.c_int, [C.c_void_p, C.POINTER(EngineLayout)]),
    "gd_batch_run": (C.c_int, [C.c_int, C.POINTER(Job), C.POINTER(BatchConfig),
                               C.POINTER(C.POINTER(ReplicaResult)), C.POINTER(McLog),
                               C.POINTER(BatchReport)]),
    "gd_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "gd_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_longlong)]),
    "gd_server_wait": (C.c_int, [C.c_void_p, C.c_longlong, C.c_double,
                                 C.POINTER(ServerJobInfo)]),
    "gd_server_fetch": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(McLog),
                                  C.POINTER(ReplicaResult), C.POINTER(ServerJobInfo)]),
    "gd_server_cancel": (C.c_int, [C.c_void_p, C.c_longlong]),
    "gd_server_drain": (C.c_int, [C.c_void_p]),
    "gd_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "gd_server_destroy": (None, [C.c_void_p]),
    "gd_optimize_enepara": (C.c_int, [C.POINTER(EnePara0), C.POINTER(EnePara)]),
    "gd_set_temperature": (C.c_int, [C.POINTER(Temp), C.c_int, C.c_float, C.c_float]),
    "gd_set_replica": (C.c_int, [C.POINTER(Replica), ComplexSize]),
    "gd_load_inputs": (C.c_int, [C.POINTER(InputFiles), C.POINTER(EnePara0),
                                 C.POINTER(C.c_void_p)]),
    "gd_inputs_get": (C.c_int, [C.c_void_p, C.POINTER(InputsView)]),
    "gd_inputs_free": (None, [C.c_void_p]),
    "gd_load_enepara": (C.c_int, [C.c_char_p, C.POINTER(EnePara0)]),
    "gd_load_weight": (C.c_int, [C.c_char_p, C.POINTER(EnePara0)]),
    "gd_load_norpara": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(EnePara0)]),
    "gd_last_error": (C.c_char_p, []),
    "gd_device_count": (C.c_int, []),
    "gd_valu_probe_kinds": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int]),
    "gd_valu_kind_name": (C.c_char_p, [C.c_int]),
    "gd_valu_probe": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "gd_version": (C.c_char_p, []),
}


def lib():
    """Load libgdremc.so (built in-tree); raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run python gpudock_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class GDError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise GDError(f"gd status {rc}: {lib().gd_last_error().decode()}")
    return rc


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _fp(a):
    return a.ctypes.data_as(FP)


def _ip(a):
    return a.ctypes.data_as(IP)


class Complex:
    """One prepared complex as numpy arrays (the inputs of Run()).

    Fields (shapes):
      lig_xyz [n_lig, lna, 3] centroid-relative conformers; lig_t, lig_c, lig_n [lna];
      lig_center [3]; prt_xyz [n_prt, pnp, 3]; prt_t, prt_c, prt_seq3r, prt_hpcflag
      [n_prt, pnp] int; prt_ele, prt_hpp [n_prt, pnp]; pocket_center [3];
      psp [24, n_res]; kde_xyz [pnk, 3]; kde_t [pnk]; mcs_xyz [pos, n_field, 3]
      (MCS_INVALID where absent); mcs_tcc [pos]; enepara (EnePara ctypes struct).
    """

    def __init__(self, **kw):
        self.lig_xyz = _f32(kw["lig_xyz"])
        self.lig_t = _i32(kw["lig_t"])
        self.lig_c = _f32(kw["lig_c"])
        self.lig_n = _i32(kw.get("lig_n", np.arange(len(self.lig_t))))
        self.lig_center = _f32(kw["lig_center"])
        self.prt_xyz = _f32(kw["prt_xyz"])
        self.prt_t = _i32(kw["prt_t"])
        self.prt_c = _i32(kw["prt_c"])
        self.prt_ele = _f32(kw["prt_ele"])
        self.prt_seq3r = _i32(kw["prt_seq3r"])
        self.prt_hpcflag = _i32(kw["prt_hpcflag"])
        self.prt_hpp = _f32(kw["prt_hpp"])
        self.pocket_center = _f32(kw["pocket_center"])
        self.psp = _f32(kw["psp"])
        self.kde_xyz = _f32(kw["kde_xyz"]).reshape(-1, 3)
        self.kde_t = _i32(kw["kde_t"])
        self.mcs_xyz = _f32(kw["mcs_xyz"])
        self.mcs_tcc = _f32(kw["mcs_tcc"])
        self.enepara = kw["enepara"]
        self._keep = []
        self._build()

    @property
    def size(self):
        s = ComplexSize()
        s.n_lig, s.lna = self.lig_xyz.shape[0], self.lig_xyz.shape[1]
        s.n_prt, s.pnp = self.prt_xyz.shape[0], self.prt_xyz.shape[1]
        s.pnk = self.kde_xyz.shape[0]
        s.pos = self.mcs_xyz.shape[0]
        s.n_tmp = 1
        s.n_rep = s.n_lig * s.n_prt
        return s

    def sized(self, n_tmp):
        s = self.size
        s.n_tmp = n_tmp
        s.n_rep = s.n_lig * s.n_prt * n_tmp
        return s

    def _build(self):
        k = self._keep
        lx = _f32(self.lig_xyz[..., 0].ravel())
        ly = _f32(self.lig_xyz[..., 1].ravel())
        lz = _f32(self.lig_xyz[..., 2].ravel())
        k += [lx, ly, lz]
        self.c_lig = Ligand(_fp(lx), _fp(ly), _fp(lz), _ip(self.lig_t), _fp(self.lig_c),
                            _ip(self.lig_n), (F * 3)(*self.lig_center))
        px = _f32(self.prt_xyz[..., 0].ravel())
        py = _f32(self.prt_xyz[..., 1].ravel())
        pz = _f32(self.prt_xyz[..., 2].ravel())
        k += [px, py, pz]
        self.c_prt = Protein(_fp(px), _fp(py), _fp(pz), _ip(self.prt_t), _ip(self.prt_c),
                             _fp(self.prt_ele), _ip(self.prt_seq3r),
                             _ip(self.prt_hpcflag), _fp(self.prt_hpp),
                             (F * 3)(*self.pocket_center))
        assert self.psp.shape[0] == MAXTP2
        self.c_psp = Psp(_fp(self.psp), self.psp.shape[1])
        kx = _f32(self.kde_xyz[:, 0])
        ky = _f32(self.kde_xyz[:, 1])
        kz = _f32(self.kde_xyz[:, 2])
        k += [kx, ky, kz]
        self.c_kde = Kde(_fp(kx), _fp(ky), _fp(kz), _ip(self.kde_t))
        nf = self.mcs_xyz.shape[1] if self.mcs_xyz.ndim == 3 else 1
        mx = _f32(self.mcs_xyz[..., 0].ravel()) if self.mcs_xyz.size else np.zeros(1, np.float32)
        my = _f32(self.mcs_xyz[..., 1].ravel()) if self.mcs_xyz.size else np.zeros(1, np.float32)
        mz = _f32(self.mcs_xyz[..., 2].ravel()) if self.mcs_xyz.size else np.zeros(1, np.float32)
        tcc = self.mcs_tcc if self.mcs_tcc.size else np.zeros(1, np.float32)
        k += [mx, my, mz, tcc]
        self.c_mcs = Mcs(_fp(mx), _fp(my), _fp(mz), _fp(tcc), nf)

    def args(self):
        return (C.byref(self.c_lig), C.byref(self.c_prt), C.byref(self.c_psp),
                C.byref(self.c_kde), C.byref(self.c_mcs), C.byref(self.enepara))


def temperatures(n_tmp, floor, ceiling):
    """SetTemperature (util.C:490-532) through the library."""
    arr = (Temp * n_tmp)()
    check(lib().gd_set_temperature(arr, n_tmp, floor, ceiling))
    return arr


def load_enepara0(ff_path, weight_path=None, nor_a=None, nor_b=None):
    ep0 = EnePara0()
    check(lib().gd_load_enepara(ff_path.encode(), C.byref(ep0)))
    if weight_path:
        check(lib().gd_load_weight(weight_path.encode(), C.byref(ep0)))
    if nor_a and nor_b:
        check(lib().gd_load_norpara(nor_a.encode(), nor_b.encode(), C.byref(ep0)))
    return ep0


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def load_complex(ligand_sdf, protein_pdb, lhm_ff, ligand_id, enepara0, molid="MOLID",
                 flags=0):
    """Read one complex from its SD, PDB and LHM files and prepare it.

    The sequence of the reference's dock.C:40-42,70-74: loadLigand,
    loadProtein, loadLHM (load.C:99-1126), then OptimizeLigand,
    OptimizeProtein, OptimizePsp, OptimizeKde and OptimizeMcs
    (util.C:199-402), all in libgdremc (gd_load_inputs).  enepara0 is the
    EnePara0 of load_enepara0(); the returned Complex carries
    optimize_enepara(enepara0).  Extra attributes: lens_rmsd [n_lig],
    prt_d and prt_r [n_prt, pnp] (residue code and index of each point),
    ligand_name.  flags: LOAD_PLATE_ALL_MODELS builds the peptide-plate
    points of every PDB model from N, C and O; by default models from the
    third on follow the reference (load.C:886), which takes C and O of
    residue serial 0 there.
    """
    files = InputFiles(os.fsencode(ligand_sdf), molid.encode() if molid else None,
                       os.fsencode(protein_pdb), os.fsencode(lhm_ff),
                       ligand_id.encode() if ligand_id is not None else None, flags)
    h = C.c_void_p()
    check(lib().gd_load_inputs(C.byref(files), C.byref(enepara0), C.byref(h)))
    try:
        v = InputsView()
        check(lib().gd_inputs_get(h, C.byref(v)))
        sz = v.size
        n_lig, lna, n_prt, pnp = sz.n_lig, sz.lna, sz.n_prt, sz.pnp
        nl, npt, nf = n_lig * lna, n_prt * pnp, v.mcs.n_field
        lig_xyz = np.stack([_copy(v.lig.x, nl, np.float32), _copy(v.lig.y, nl, np.float32),
                            _copy(v.lig.z, nl, np.float32)], -1).reshape(n_lig, lna, 3)
        prt_xyz = np.stack([_copy(v.prt.x, npt, np.float32), _copy(v.prt.y, npt, np.float32),
                            _copy(v.prt.z, npt, np.float32)], -1).reshape(n_prt, pnp, 3)
        pnk, pos = sz.pnk, sz.pos
        kde_xyz = np.stack([_copy(v.kde.x, pnk, np.float32), _copy(v.kde.y, pnk, np.float32),
                            _copy(v.kde.z, pnk, np.float32)], -1).reshape(pnk, 3)
        mcs_xyz = np.stack([_copy(v.mcs.x, pos * nf, np.float32),
                            _copy(v.mcs.y, pos * nf, np.float32),
                            _copy(v.mcs.z, pos * nf, np.float32)], -1).reshape(pos, nf, 3)
        shp = (n_prt, pnp)
        cx = Complex(
            lig_xyz=lig_xyz, lig_t=_copy(v.lig.t, lna, np.int32),
            lig_c=_copy(v.lig.c, lna, np.float32), lig_n=_copy(v.lig.n, lna, np.int32),
            lig_center=np.array(v.lig.center[:], np.float32), prt_xyz=prt_xyz,
            prt_t=_copy(v.prt.t, npt, np.int32).reshape(shp),
            prt_c=_copy(v.prt.c, npt, np.int32).reshape(shp),
            prt_ele=_copy(v.prt.ele, npt, np.float32).reshape(shp),
            prt_seq3r=_copy(v.prt.seq3r, npt, np.int32).reshape(shp),
            prt_hpcflag=_copy(v.prt.c0_and_d12_or_c2, npt, np.int32).reshape(shp),
            prt_hpp=_copy(v.prt.hpp, npt, np.float32).reshape(shp),
            pocket_center=np.array(v.prt.pocket_center[:], np.float32),
            psp=_copy(v.psp.psp, MAXTP2 * v.psp.n_res, np.float32).reshape(MAXTP2,
                                                                             v.psp.n_res),
            kde_xyz=kde_xyz, kde_t=_copy(v.kde.t, pnk, np.int32), mcs_xyz=mcs_xyz,
            mcs_tcc=_copy(v.mcs.tcc, pos, np.float32), enepara=optimize_enepara(enepara0))
        cx.lens_rmsd = _copy(v.lens_rmsd, n_lig, np.float32)
        cx.prt_d = _copy(v.prt_d, npt, np.int32).reshape(shp)
        cx.prt_r = _copy(v.prt_r, npt, np.int32).reshape(shp)
        cx.ligand_name = (v.ligand_name or b"").decode(errors="replace")
        return cx
    finally:
        lib().gd_inputs_free(h)


def optimize_enepara(ep0):
    ep = EnePara()
    check(lib().gd_optimize_enepara(C.byref(ep0), C.byref(ep)))
    return ep


def default_options(seed=1, complex_id=0, exchange=1, away=0, term_order=0, device=0,
                    record_stride=0, record_prefix=None, mcc=0, combine=0, fixed_draws=0,
                    reverse=0):
    return Options(seed, complex_id, exchange, away, term_order, device, record_stride,
                   record_prefix.encode() if isinstance(record_prefix, str) else record_prefix,
                   mcc, combine, fixed_draws, reverse)


def mcpara(steps_total, steps_per_exchange, move_scale, steps_per_dump=None):
    """McPara.  The run follows the reference's schedule (gpusingle.cu:8-20):
    windows of steps_per_dump steps, each of whole launches of
    steps_per_exchange steps with the exchange mode restarting at 0;
    steps_per_dump None = steps_total (one window)."""
    if steps_per_dump is None:
        steps_per_dump = max(int(steps_total), 1)
    return McPara(steps_total, steps_per_dump, steps_per_exchange, (F * 6)(*move_scale))


def results_to_numpy(res):
    n = len(res)
    out = {
        "move_final": np.array([list(r.move_final) for r in res], np.float32).reshape(n, 6),
        "energy_final": np.array([list(r.energy_final) for r in res], np.float32).reshape(n, 10),
        "move_best": np.array([list(r.move_best) for r in res], np.float32).reshape(n, 6),
        "energy_best": np.array([list(r.energy_best) for r in res], np.float32).reshape(n, 10),
        "energy_best_fp64": np.array([list(r.energy_best_fp64) for r in res]).reshape(n, 10),
        "step_best": np.array([r.step_best for r in res], np.int32),
        "idx_tmp_final": np.array([r.idx_tmp_final for r in res], np.int32),
        "acs_mc": np.array([r.acs_mc for r in res], np.int32),
        "acs_exchg": np.array([r.acs_exchg for r in res], np.int32),
        "cmcc_final": np.array([r.cmcc_final for r in res], np.float32),
        "cmcc_best": np.array([r.cmcc_best for r in res], np.float32),
    }
    return out


def replica_table(size):
    """SetReplica (util.C:537-558) through the library: [n_rep] Replica."""
    arr = (Replica * size.n_rep)()
    check(lib().gd_set_replica(arr, size))
    return arr


def run(cx, temps, mc, opt=None, n_tmp=None, replica=None):
    """gd_run (drop-in for Run()); returns (results dict, McLog).  replica:
    a Replica table (None = the SetReplica layout)."""
    n_tmp = n_tmp or len(temps)
    size = cx.sized(n_tmp)
    res = (ReplicaResult * size.n_rep)()
    acs = np.zeros(size.n_rep, np.int32)
    acx = np.zeros(size.n_rep, np.int32)
    log = McLog(0, 0, 0, 0, _ip(acs), 0, _ip(acx))
    opt = opt or default_options()
    check(lib().gd_run(*cx.args(), temps, replica, C.byref(mc), C.byref(log), size,
                       C.byref(opt), res))
    return results_to_numpy(res), log


def valu_probe(waves_per_simd=8, device=0):
    """gd_valu_probe: (ns per wavefront FMA instruction per SIMD, device FMA
    TFLOP/s, SIMDs of the device) with every operand in vector registers."""
    ns, tf, n = C.c_double(), C.c_double(), C.c_int()
    check(lib().gd_valu_probe(device, waves_per_simd, C.byref(ns), C.byref(tf), C.byref(n)))
    return ns.value, tf.value, n.value


VALU_KINDS = 10


def valu_probe_kinds(waves_per_simd=8, device=0):
    """gd_valu_probe_kinds: {instruction kind: ns per wavefront instruction
    per SIMD} for the kinds the energy loops are made of."""
    ns = (C.c_double * VALU_KINDS)()
    check(lib().gd_valu_probe_kinds(device, waves_per_simd, ns, VALU_KINDS))
    return {lib().gd_valu_kind_name(k).decode(): ns[k] for k in range(VALU_KINDS)}


def score_poses(cx, lig_conf, prt_conf, moves, term_order=0, device=0, fp64=True,
                noalloc=False):
    """gd_score_poses: FP32 canonical energies (and FP64) of poses, slot order.
    noalloc: through gd_score_poses_noalloc (static device workspace)."""
    lig_conf = _i32(lig_conf)
    prt_conf = _i32(prt_conf)
    moves = _f32(moves).reshape(-1, 6)
    n = len(lig_conf)
    e = np.zeros((n, 10), np.float32)
    e64 = np.zeros((n, 10), np.float64) if fp64 else None
    fn = lib().gd_score_poses_noalloc if noalloc else lib().gd_score_poses
    check(fn(*cx.args(), cx.size, n, _ip(lig_conf), _ip(prt_conf), _fp(moves), term_order,
             device, _fp(e), e64.ctypes.data_as(C.POINTER(C.c_double)) if fp64 else None))
    return (e, e64) if fp64 else e


class Session:
    """gd_session_*: complex and replica state resident on the GPU."""

    def __init__(self, cx, temps, mc, opt=None):
        self.cx, self.temps, self.mc = cx, temps, mc
        self.opt = opt or default_options()
        self.size = cx.sized(len(temps))
        h = C.c_void_p()
        check(lib().gd_session_create(*cx.args(), temps, None, C.byref(mc), self.size,
                                      C.byref(self.opt), C.byref(h)))
        self.h = h

    def init(self):
        check(lib().gd_session_init(self.h))

    def cycles(self, n):
        check(lib().gd_session_cycles(self.h, n))

    def sync(self):
        check(lib().gd_session_sync(self.h))

    def steps_done(self):
        """MC steps queued so far (the schedule's count, gpusingle.cu:8-20)."""
        return int(lib().gd_session_steps_done(self.h))

    def set_timing(self, on):
        check(lib().gd_session_set_timing(self.h, int(on)))

    def timing(self):
        a, b = C.c_double(), C.c_int()
        c, d = C.c_double(), C.c_int()
        check(lib().gd_session_timing(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return a.value, b.value, c.value, d.value

    def layout(self):
        v = [C.c_int() for _ in range(4)]
        check(lib().gd_session_layout(self.h, *[C.byref(x) for x in v]))
        return {"lds_bytes": v[0].value, "kde_in_lds": v[1].value,
                "n_tasks": v[2].value, "waves_per_group": v[3].value}

    def results(self):
        res = (ReplicaResult * self.size.n_rep)()
        log = McLog()
        check(lib().gd_session_results(self.h, C.byref(log), res))
        return results_to_numpy(res), log

    def records(self):
        """(first step, [n_rep, n_step] records of STEP_RECORD_DTYPE) of the
        current steps_per_dump window."""
        first, n = C.c_int(), C.c_int()
        check(lib().gd_session_records(self.h, C.byref(first), C.byref(n), None))
        out = np.zeros((self.size.n_rep, n.value), STEP_RECORD_DTYPE)
        check(lib().gd_session_records(self.h, C.byref(first), C.byref(n),
                                       out.ctypes.data_as(C.c_void_p)))
        return first.value, out

    def close(self):
        if self.h:
            lib().gd_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _struct_dict(st):
    out = {}
    for name, typ in st._fields_:
        v = getattr(st, name)
        if isinstance(v, C.Structure):
            v = _struct_dict(v)
        elif isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


class JobSpec:
    """One gd_job: a complex, its temperature ladder, McPara and options.

    Holds every ctypes object the job's pointers refer to."""

    def __init__(self, cx, temps, mc, opt=None, replica=None):
        self.cx, self.temps, self.mc = cx, temps, mc
        self.opt = opt or default_options()
        self.size = cx.sized(len(temps))
        self.replica = replica
        rep = C.cast(replica, C.POINTER(Replica)) if replica is not None else None
        self.c = Job(C.pointer(cx.c_lig), C.pointer(cx.c_prt), C.pointer(cx.c_psp),
                     C.pointer(cx.c_kde), C.pointer(cx.c_mcs), C.pointer(cx.enepara),
                     C.cast(temps, C.POINTER(Temp)), rep, C.pointer(mc), self.size,
                     C.pointer(self.opt))

    @property
    def n_rep(self):
        return self.size.n_rep

    def footprint(self, device=0):
        b = C.c_size_t()
        check(lib().gd_job_footprint(C.byref(self.c), device, C.byref(b)))
        return b.value


class Engine:
    """gd_engine_*: many complexes resident on one GPU, advanced together."""

    def __init__(self, device=0, arena_bytes=1 << 30, wpg=8, persistent=0, use_graph=1,
                 kde_policy=0, occ=0, split=0):
        cfg = EngineConfig(device, arena_bytes, wpg, persistent, use_graph, kde_policy, 0, occ,
                           split)
        h = C.c_void_p()
        check(lib().gd_engine_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.jobs = {}

    def add(self, job):
        hd = C.c_int()
        check(lib().gd_engine_add(self.h, C.byref(job.c), C.byref(hd)))
        self.jobs[hd.value] = job
        return hd.value

    def remove(self, handle):
        check(lib().gd_engine_remove(self.h, handle))
        self.jobs.pop(handle, None)

    def cycles(self, n):
        check(lib().gd_engine_cycles(self.h, n))

    def sync(self):
        check(lib().gd_engine_sync(self.h))

    def cycles_left(self, handle):
        rc = lib().gd_engine_cycles_left(self.h, handle)
        if rc < 0:
            check(rc)
        return rc

    def results(self, handle):
        job = self.jobs[handle]
        res = (ReplicaResult * job.n_rep)()
        log = McLog()
        check(lib().gd_engine_results(self.h, handle, C.byref(log), res))
        return results_to_numpy(res), log

    def autotune(self):
        rep = TuneReport()
        check(lib().gd_engine_autotune(self.h, C.byref(rep)))
        return _struct_dict(rep)

    def set_timing(self, on):
        check(lib().gd_engine_set_timing(self.h, int(on)))

    def timing(self):
        a, b = C.c_double(), C.c_int()
        c, d = C.c_double(), C.c_int()
        check(lib().gd_engine_timing(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return a.value, b.value, c.value, d.value

    def layout(self):
        info = EngineLayout()
        check(lib().gd_engine_layout(self.h, C.byref(info)))
        return _struct_dict(info)

    def close(self):
        if self.h:
            lib().gd_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run(jobs, device=0, wpg=0, persistent=0, no_graph=0, autotune=0, kde_policy=0,
              occ=0, split=0):
    """gd_batch_run over JobSpecs; returns ([results dict per job], report dict)."""
    n = len(jobs)
    arr = (Job * n)(*[j.c for j in jobs])
    res = [(ReplicaResult * j.n_rep)() for j in jobs]
    ptrs = (C.POINTER(ReplicaResult) * n)(*[C.cast(r, C.POINTER(ReplicaResult)) for r in res])
    logs = (McLog * n)()
    rep = BatchReport()
    cfg = BatchConfig(device, wpg, persistent, no_graph, autotune, kde_policy, occ, split)
    check(lib().gd_batch_run(n, arr, C.byref(cfg), ptrs, logs, C.byref(rep)))
    return [results_to_numpy(r) for r in res], _struct_dict(rep)


class Server:
    """gd_server_*: request loop with continuous batching on one GPU."""

    def __init__(self, device=0, arena_bytes=0, wpg=8, persistent=0, use_graph=0,
                 kde_policy=0, max_resident_replicas=0, cycles_per_tick=1,
                 max_ticks_in_flight=2, keep_results=1, readback_bytes=0, occ=0, split=0,
                 lane_max_pairs=(), lane_split=(), lane_max_jobs=(), lane_prio=(),
                 lane_max_total=()):
        """lane_max_pairs: upper bounds of lna x pnp for all lanes but the
        last (len + 1 lanes; empty = one engine); lane_split: wavefronts per
        replica of each lane (0 = as split); lane_max_jobs: jobs a lane holds
        queued or resident before it passes new ones on (0 = no limit);
        lane_prio: issue priority (0-3) of each lane's wavefronts;
        lane_max_total: jobs in the whole server above which a lane passes
        new ones on (0 = no limit)."""
        bounds = list(lane_max_pairs)
        if (len(bounds) > SERVER_MAX_LANES - 1 or len(lane_split) > SERVER_MAX_LANES
                or len(lane_max_jobs) > SERVER_MAX_LANES or len(lane_prio) > SERVER_MAX_LANES
                or len(lane_max_total) > SERVER_MAX_LANES):
            raise GDError(f"at most {SERVER_MAX_LANES} lanes")
        cfg = ServerConfig(device, arena_bytes, wpg, persistent, use_graph, kde_policy,
                           max_resident_replicas, cycles_per_tick, max_ticks_in_flight,
                           keep_results, readback_bytes, occ, split, len(bounds) + 1,
                           (C.c_int * (SERVER_MAX_LANES - 1))(*bounds),
                           (C.c_int * SERVER_MAX_LANES)(*lane_split),
                           (C.c_int * SERVER_MAX_LANES)(*lane_max_jobs),
                           (C.c_int * SERVER_MAX_LANES)(*lane_prio),
                           (C.c_int * SERVER_MAX_LANES)(*lane_max_total))
        h = C.c_void_p()
        check(lib().gd_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.n_rep = {}

    def submit(self, job):
        t = C.c_longlong()
        check(lib().gd_server_submit(self.h, C.byref(job.c), C.byref(t)))
        self.n_rep[t.value] = job.n_rep
        return t.value

    def wait(self, ticket, timeout=-1.0):
        info = ServerJobInfo()
        check(lib().gd_server_wait(self.h, ticket, timeout, C.byref(info)))
        return _struct_dict(info)

    def fetch(self, ticket):
        n = self.n_rep.pop(ticket)
        res = (ReplicaResult * n)()
        log = McLog()
        info = ServerJobInfo()
        check(lib().gd_server_fetch(self.h, ticket, C.byref(log), res, C.byref(info)))
        return results_to_numpy(res), log, _struct_dict(info)

    def cancel(self, ticket):
        check(lib().gd_server_cancel(self.h, ticket))

    def drain(self):
        check(lib().gd_server_drain(self.h))

    def stats(self):
        st = ServerStats()
        check(lib().gd_server_get_stats(self.h, C.byref(st)))
        return _struct_dict(st)

    def close(self):
        if self.h:
            lib().gd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
