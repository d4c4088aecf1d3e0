"""MI355X HDNNP engine: Python host mirror of the reference potential API.

Binds lib/libhdnnp_amd.so (include/hdnnp_capi.h) with ctypes. The names
mirror the reference Fortran interface (src/potentials/hdnnps.f90):

    h = parseRuNNerInput(path, precision="fp64")        # hdnnps.f90:34
    E, F, dEdLat = hdnnpEnergyAndForces(h, ats)          # hdnnps.f90:333
    stress = dEdLat2Stress(lat, dEdLat)                  # atomicStructure.f90:679

Errors raise HdnnpError (status code + message) instead of the reference's
print + stop. There is no CPU fallback: without the built library every call
raises.
"""
from __future__ import annotations

import ctypes
import sys
import os

import numpy as np

from .workloads import Structure  # noqa: F401  (host mirror of atStruct)
from . import readers  # noqa: F401  (structure file formats of atomicStructure.f90)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhdnnp_amd.so")

FP64, MIXED, MIXED_BF16 = 0, 1, 2
_PREC = {"fp64": FP64, "mixed": MIXED, "mixed-bf16": MIXED_BF16, FP64: FP64, MIXED: MIXED, MIXED_BF16: MIXED_BF16}
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

EXPORTS = ("hdnnp_create", "hdnnp_evaluate", "hdnnp_evaluate_batch", "hdnnp_evaluate_batch_device", "hdnnp_evaluate_device",
           "hdnnp_evaluate_batch_device_noalloc",
           "hdnnp_stage_neighbor_lists", "hdnnp_stage_get_neighbor_lists", "hdnnp_stage_symfunctions",
           "hdnnp_stage_nns",
           "hdnnp_create_committee", "hdnnp_committee_size", "hdnnp_last_committee",
           "hdnnp_md_create", "hdnnp_md_step", "hdnnp_md_get", "hdnnp_md_stats", "hdnnp_md_counters", "hdnnp_md_destroy", "hdnnp_neighbor_count",
           "hdnnp_get_neighbors", "hdnnp_get_symfunctions", "hdnnp_get_model_info",
           "hdnnp_get_sf_table", "hdnnp_get_weights", "hdnnp_dEdLat_to_stress",
           "hdnnp_last_error", "hdnnp_last_timings", "hdnnp_set_stage_timings", "hdnnp_set_graphs", "hdnnp_graph_counts", "hdnnp_destroy",
           "hdnnp_set_tile", "hdnnp_get_tile", "hdnnp_autotune",
           "hdnnp_server_create", "hdnnp_server_submit", "hdnnp_server_wait", "hdnnp_server_cancel",
           "hdnnp_server_md_open",
           "hdnnp_server_md_step", "hdnnp_server_md_wait", "hdnnp_server_md_get", "hdnnp_server_md_close",
           "hdnnp_server_stats", "hdnnp_server_last_error", "hdnnp_server_destroy")

SERVER_STATS = ("evaluations", "eval_atoms", "md_requests", "md_atom_steps", "eval_p50_ms", "eval_p99_ms",
                "md_p50_ms", "md_p99_ms", "batches", "mean_batch", "peak_reserved_bytes", "budget_bytes",
                "seconds", "max_busy_lanes", "refused", "md_steps")
HDNNP_ERR_TIMEOUT = 7
HDNNP_ERR_CANCELLED = 8


class HdnnpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ModelInfo(ctypes.Structure):
    _fields_ = [("nelements", ctypes.c_int32), ("Z", ctypes.c_int32 * 16), ("nsf", ctypes.c_int32 * 16),
                ("nlayers", ctypes.c_int32), ("nodes", ctypes.c_int32 * 8), ("act", ctypes.c_char * 8),
                ("scalesf", ctypes.c_int32), ("centersf", ctypes.c_int32), ("maxrc", ctypes.c_double),
                ("atom_energy", ctypes.c_double * 16), ("emin", ctypes.c_double),
                ("emax", ctypes.c_double), ("cutoff_type", ctypes.c_int32), ("scale_range", ctypes.c_int32),
                ("scale_min", ctypes.c_double), ("scale_max", ctypes.c_double)]


_LIB = None


def load_library():
    """Load lib/libhdnnp_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HdnnpError(5, f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C "
                                f"{HERE})")
        # torch bundles its own HIP runtime; two runtimes share the device in
        # one process only when torch's initialises first, so bring it up
        # before the library if torch is in use here
        if "torch" in sys.modules:
            sys.modules["torch"].cuda.is_available()
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        lib.hdnnp_create.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(vp)]
        lib.hdnnp_evaluate.argtypes = [vp, ctypes.c_int32, _ip, _dp, ctypes.c_int32, _dp, _dp, _dp, _dp]
        lib.hdnnp_evaluate_device.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.c_int32, _dp, _dp, vp, _dp, vp]
        lib.hdnnp_evaluate_batch.argtypes = [vp, ctypes.c_int32, _ip, _ip, _dp, _ip, _dp, _dp, _dp, _dp]
        lib.hdnnp_md_create.argtypes = [vp, ctypes.c_int32, _ip, _ip, _dp, _dp, _dp, _ip, _dp, ctypes.c_double,
                                        ctypes.c_double, ctypes.POINTER(vp)]
        lib.hdnnp_md_step.argtypes = [vp, ctypes.c_int32, _dp, _dp]
        lib.hdnnp_md_get.argtypes = [vp, _dp, _dp, _dp]
        lib.hdnnp_md_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        lib.hdnnp_md_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        lib.hdnnp_md_destroy.argtypes = [vp]
        lib.hdnnp_md_destroy.restype = None
        lib.hdnnp_create_committee.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32,
                                               ctypes.POINTER(vp)]
        lib.hdnnp_committee_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        lib.hdnnp_last_committee.argtypes = [vp, ctypes.c_int32, _dp, _dp]
        lib.hdnnp_evaluate_batch_device.argtypes = [vp, ctypes.c_int32, _ip, vp, vp, _ip, _dp, _dp, vp, _dp, vp]
        lib.hdnnp_evaluate_batch_device_noalloc.argtypes = [vp, ctypes.c_int32, _ip, vp, vp, _ip, _dp, _dp, vp, _dp,
                                                            vp]
        i64p_ = ctypes.POINTER(ctypes.c_int64)
        lib.hdnnp_stage_neighbor_lists.argtypes = [vp, ctypes.c_int32, _ip, _dp, ctypes.c_int32, _dp, i64p_, _ip]
        lib.hdnnp_stage_get_neighbor_lists.argtypes = [vp, ctypes.c_int64, _ip, _dp]
        lib.hdnnp_stage_symfunctions.argtypes = [vp, i64p_, ctypes.c_int32, _ip, _dp, ctypes.c_int32, _dp,
                                                 ctypes.c_int32, ctypes.c_int32, _dp, _dp]
        lib.hdnnp_stage_nns.argtypes = [vp, vp, i64p_, ctypes.c_int32, _ip, _dp, ctypes.c_int32, _dp,
                                        ctypes.c_int32, _dp, _dp, _dp, _dp]
        lib.hdnnp_neighbor_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        lib.hdnnp_get_neighbors.argtypes = [vp, ctypes.c_int64, _ip]
        lib.hdnnp_get_symfunctions.argtypes = [vp, ctypes.c_int32, _dp, _dp, _dp]
        lib.hdnnp_get_model_info.argtypes = [vp, ctypes.POINTER(ModelInfo)]
        lib.hdnnp_get_sf_table.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _dp]
        lib.hdnnp_get_weights.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, _dp, ctypes.POINTER(ctypes.c_int64)]
        lib.hdnnp_dEdLat_to_stress.argtypes = [_dp, _dp, _dp]
        lib.hdnnp_last_error.argtypes = [vp]
        lib.hdnnp_last_error.restype = ctypes.c_char_p
        lib.hdnnp_last_timings.argtypes = [vp, _dp]
        lib.hdnnp_set_stage_timings.argtypes = [vp, ctypes.c_int32]
        lib.hdnnp_set_graphs.argtypes = [vp, ctypes.c_int32]
        lib.hdnnp_graph_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64)]
        lib.hdnnp_destroy.argtypes = [vp]
        lib.hdnnp_destroy.restype = None
        lib.hdnnp_set_tile.argtypes = [vp, ctypes.c_int32]
        lib.hdnnp_get_tile.argtypes = [vp, _ip]
        lib.hdnnp_autotune.argtypes = [vp, ctypes.c_int32, _ip, _ip, _dp, _ip, _dp, _ip, _dp]
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.hdnnp_server_create.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(vp)]
        lib.hdnnp_server_submit.argtypes = [vp, ctypes.c_int32, _ip, _dp, ctypes.c_int32, _dp, i64p]
        lib.hdnnp_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]
        lib.hdnnp_server_cancel.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
        lib.hdnnp_server_md_open.argtypes = [vp, ctypes.c_int32, _ip, _dp, _dp, _dp, ctypes.c_int32, _dp,
                                             ctypes.c_double, ctypes.c_double, i64p]
        lib.hdnnp_server_md_step.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, i64p]
        lib.hdnnp_server_md_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double, _dp, _dp, _dp]
        lib.hdnnp_server_md_get.argtypes = [vp, ctypes.c_int64, _dp, _dp, _dp]
        lib.hdnnp_server_md_close.argtypes = [vp, ctypes.c_int64]
        lib.hdnnp_server_stats.argtypes = [vp, ctypes.c_int32, _dp]
        lib.hdnnp_server_last_error.argtypes = []
        lib.hdnnp_server_last_error.restype = ctypes.c_char_p
        lib.hdnnp_server_destroy.argtypes = [vp]
        lib.hdnnp_server_destroy.restype = None
        _LIB = lib
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _check_device_inputs(torch, nat, Z, pos, force):
    """The device entry points read Z as int32[nat] and pos / write force as
    float64[nat][3] at their data pointers: anything else would be misread or
    overrun, so it is refused here, before any launch."""
    for name, t in (("Z", Z), ("pos", pos), ("force", force)):
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name} must be a torch tensor")
    if Z.dtype != torch.int32:
        raise ValueError(f"Z must be int32 (got {Z.dtype})")
    for name, t in (("pos", pos), ("force", force)):
        if t.dtype != torch.float64:
            raise ValueError(f"{name} must be float64 (got {t.dtype})")
    if tuple(Z.shape) != (nat,):
        raise ValueError(f"Z has shape {tuple(Z.shape)}, expected ({nat},)")
    for name, t in (("pos", pos), ("force", force)):
        if tuple(t.shape) != (nat, 3):
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected ({nat}, 3)")
    for name, t in (("Z", Z), ("pos", pos), ("force", force)):
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    for name, t in (("Z", Z), ("pos", pos), ("force", force)):
        if t.device.type != "cuda":
            raise ValueError(f"{name} must be in GPU memory (got {t.device})")
    if not (Z.device == pos.device == force.device):
        raise ValueError(f"Z, pos and force are on different devices ({Z.device}, {pos.device}, {force.device})")


class Hdnnp:
    """Handle of one potential (the reference's type(hdnnp), hdnnps.f90:13-21)."""

    def __init__(self, param_dir, precision="fp64"):
        """param_dir: one parameter directory, or a list of them for a committee
        (globalHdnnps.f90: the first one's symmetry functions, every member's
        networks; evaluations return the committee mean)."""
        self.lib = load_library()
        self.ctx = ctypes.c_void_p()
        self.precision = _PREC[precision]
        if isinstance(param_dir, (list, tuple)):
            dirs = (ctypes.c_char_p * len(param_dir))(*[os.fsencode(d) for d in param_dir])
            rc = self.lib.hdnnp_create_committee(len(param_dir), dirs, self.precision, ctypes.byref(self.ctx))
        else:
            rc = self.lib.hdnnp_create(os.fsencode(param_dir), self.precision, ctypes.byref(self.ctx))
        if rc:
            raise HdnnpError(rc, self.lib.hdnnp_last_error(None).decode())
        info = ModelInfo()
        self._check(self.lib.hdnnp_get_model_info(self.ctx, ctypes.byref(info)))
        self.info = info
        self.maxsf = max(info.nsf[e] for e in range(info.nelements))

    @property
    def committee_size(self) -> int:
        n = ctypes.c_int32()
        self._check(self.lib.hdnnp_committee_size(self.ctx, ctypes.byref(n)))
        return n.value

    def committee(self, nstruct: int = 1):
        """Of the last evaluation: (std of the member energies [nstruct],
        member energies [nstruct, N])."""
        sd = np.zeros(nstruct)
        me = np.zeros((nstruct, self.committee_size))
        self._check(self.lib.hdnnp_last_committee(self.ctx, nstruct, _p(sd, _dp), _p(me, _dp)))
        return sd, me

    def _check(self, rc):
        if rc:
            raise HdnnpError(rc, self.lib.hdnnp_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.hdnnp_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- model introspection -------------------------------------------------
    @property
    def elements(self):
        return [self.info.Z[e] for e in range(self.info.nelements)]

    @property
    def maxrc(self) -> float:
        return self.info.maxrc

    def sf_table(self, e: int) -> np.ndarray:
        n = self.info.nsf[e]
        rows = np.zeros((max(n, 1), 11))
        self._check(self.lib.hdnnp_get_sf_table(self.ctx, e, n, _p(rows, _dp)))
        return rows[:n]

    def weights(self, e: int) -> np.ndarray:
        cnt = ctypes.c_int64()
        self._check(self.lib.hdnnp_get_weights(self.ctx, e, 0, None, ctypes.byref(cnt)))
        w = np.zeros(cnt.value)
        self._check(self.lib.hdnnp_get_weights(self.ctx, e, cnt.value, _p(w, _dp), ctypes.byref(cnt)))
        return w

    # --- evaluation ----------------------------------------------------------
    def energy_and_forces(self, s, want_dEdLat: bool = True):
        """hdnnpEnergyAndForces (hdnnps.f90:333-371): returns (E, F[nat,3], dEdLat[3,3]|None).
        dEdLat uses the reference memory order (row k = d E / d lattice vector k)."""
        pos = np.ascontiguousarray(s.pos, dtype=np.float64)
        Z = np.ascontiguousarray(s.Z, dtype=np.int32)
        lat = np.ascontiguousarray(s.lat, dtype=np.float64)
        nat = Z.shape[0]
        E = ctypes.c_double()
        F = np.zeros((nat, 3))
        L = np.zeros((3, 3)) if want_dEdLat else None
        self._check(self.lib.hdnnp_evaluate(self.ctx, nat, _p(Z, _ip), _p(pos, _dp), int(bool(s.periodic)),
                                            _p(lat, _dp), ctypes.byref(E), _p(F, _dp), _p(L, _dp)))
        return E.value, F, L

    @property
    def tile(self) -> int:
        """atoms per wavefront of the symmetry-function kernels (1, 2 or 4)"""
        t = ctypes.c_int32()
        self._check(self.lib.hdnnp_get_tile(self.ctx, ctypes.byref(t)))
        return t.value

    @tile.setter
    def tile(self, atoms_per_wave: int):
        self._check(self.lib.hdnnp_set_tile(self.ctx, int(atoms_per_wave)))

    @property
    def graphs(self) -> bool:
        """graph replay of small evaluations (hdnnp_set_graphs)"""
        on = ctypes.c_int32()
        self._check(self.lib.hdnnp_graph_counts(self.ctx, ctypes.byref(on), None, None))
        return bool(on.value)

    @graphs.setter
    def graphs(self, on: bool):
        self._check(self.lib.hdnnp_set_graphs(self.ctx, 1 if on else 0))

    def graph_counts(self):
        """(captured, replayed) evaluations of this context"""
        c, r = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.hdnnp_graph_counts(self.ctx, None, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def autotune(self, structures):
        """hdnnp_autotune on a sample: (chosen tile, {tile: ms of the SF passes})."""
        nat = np.array([int(np.asarray(s.Z).shape[0]) for s in structures], dtype=np.int32)
        Z = np.ascontiguousarray(np.concatenate([np.asarray(s.Z, dtype=np.int32) for s in structures]))
        pos = np.ascontiguousarray(np.concatenate([np.asarray(s.pos, dtype=np.float64).reshape(-1, 3)
                                                   for s in structures]))
        per = np.array([int(bool(s.periodic)) for s in structures], dtype=np.int32)
        lat = np.ascontiguousarray(np.stack([np.asarray(s.lat, dtype=np.float64).reshape(3, 3) for s in structures]))
        t = ctypes.c_int32()
        ms = np.zeros(3)
        self._check(self.lib.hdnnp_autotune(self.ctx, len(structures), _p(nat, _ip), _p(Z, _ip), _p(pos, _dp),
                                            _p(per, _ip), _p(lat, _dp), ctypes.byref(t), _p(ms, _dp)))
        return t.value, dict(zip((1, 2, 4), ms.tolist()))

    def energy_and_forces_batch(self, structures, want_dEdLat: bool = True):
        """hdnnp_evaluate_batch: many structures in one evaluation. Returns
        (E[nstruct], [F_k (nat_k, 3)], dEdLat[nstruct, 3, 3] | None)."""
        ns = len(structures)
        nat = np.array([int(np.asarray(s.Z).shape[0]) for s in structures], dtype=np.int32)
        Z = np.ascontiguousarray(np.concatenate([np.asarray(s.Z, dtype=np.int32) for s in structures])
                                 if ns else np.zeros(0, np.int32))
        pos = np.ascontiguousarray(np.concatenate([np.asarray(s.pos, dtype=np.float64).reshape(-1, 3)
                                                   for s in structures]) if ns else np.zeros((0, 3)))
        per = np.array([int(bool(s.periodic)) for s in structures], dtype=np.int32)
        lat = np.ascontiguousarray(np.stack([np.asarray(s.lat, dtype=np.float64).reshape(3, 3)
                                             for s in structures]) if ns else np.zeros((0, 3, 3)))
        E = np.zeros(ns)
        F = np.zeros((int(nat.sum()), 3))
        L = np.zeros((ns, 3, 3)) if want_dEdLat else None
        self._check(self.lib.hdnnp_evaluate_batch(self.ctx, ns, _p(nat, _ip), _p(Z, _ip), _p(pos, _dp),
                                                  _p(per, _ip), _p(lat, _dp), _p(E, _dp), _p(F, _dp),
                                                  _p(L, _dp)))
        off = np.concatenate([[0], np.cumsum(nat)])
        return E, [F[off[k]:off[k + 1]] for k in range(ns)], L

    def evaluate_batch_torch(self, nat, Z, pos, periodic, lat, force=None, stream=None, noalloc=False):
        """hdnnp_evaluate_batch_device: nat/periodic (nstruct,), lat (nstruct, 3, 3)
        host arrays; Z (int32 [sum nat]), pos and force (float64 [sum nat, 3])
        contiguous torch tensors on one GPU. Returns (E[nstruct], force tensor,
        dEdLat[nstruct, 3, 3]). Raises ValueError for any other input before
        anything is launched. noalloc: through
        hdnnp_evaluate_batch_device_noalloc (no device allocation; pass
        `force` to allocate nothing here either)."""
        import torch
        nat = np.ascontiguousarray(nat, dtype=np.int32).reshape(-1)
        per = np.ascontiguousarray(periodic, dtype=np.int32).reshape(-1)
        ns = nat.shape[0]
        if per.shape[0] != ns:
            raise ValueError(f"periodic has {per.shape[0]} entries for {ns} structures")
        if (nat < 0).any():
            raise ValueError("negative atom count")
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        if lat.size != 9 * ns:
            raise ValueError(f"lat must hold 9 values per structure ({9 * ns}), got {lat.size}")
        total = int(nat.sum())
        if force is None:
            force = torch.empty((total, 3), dtype=torch.float64, device=getattr(pos, "device", None))
        _check_device_inputs(torch, total, Z, pos, force)
        E = np.zeros(ns)
        L = np.zeros((ns, 3, 3))
        entry = self.lib.hdnnp_evaluate_batch_device_noalloc if noalloc else self.lib.hdnnp_evaluate_batch_device
        self._check(entry(
            self.ctx, ns, _p(nat, _ip), ctypes.c_void_p(Z.data_ptr()), ctypes.c_void_p(pos.data_ptr()),
            _p(per, _ip), _p(lat, _dp), _p(E, _dp), ctypes.c_void_p(force.data_ptr()), _p(L, _dp),
            ctypes.c_void_p(stream or 0)))
        return E, force, L

    def evaluate_device(self, nat: int, Z_dev, pos_dev, force_dev, periodic: bool, lat, stream=None):
        """hdnnp_evaluate_device on raw device pointers (ints): Z_dev int32[nat],
        pos_dev and force_dev float64[nat][3], all in device memory; lat a (3,3)
        host array. The caller guarantees the buffers hold nat atoms. Returns
        (E, dEdLat)."""
        nat = int(nat)
        if nat < 0:
            raise ValueError("nat < 0")
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        if lat.size != 9:
            raise ValueError("lat must hold 9 values")
        E = ctypes.c_double()
        L = np.zeros((3, 3))
        self._check(self.lib.hdnnp_evaluate_device(self.ctx, nat, ctypes.c_void_p(Z_dev), ctypes.c_void_p(pos_dev),
                                                   int(bool(periodic)), _p(lat, _dp), ctypes.byref(E),
                                                   ctypes.c_void_p(force_dev), _p(L, _dp),
                                                   ctypes.c_void_p(stream or 0)))
        return E.value, L

    def evaluate_torch(self, Z, pos, periodic: bool, lat, force=None, stream=None):
        """Z (int32 [nat]), pos and force (float64 [nat, 3]) contiguous torch
        tensors on one GPU; returns (E, force, dEdLat). Raises ValueError for
        any other input before anything is launched."""
        import torch
        if not isinstance(Z, torch.Tensor) or Z.dim() != 1:
            raise ValueError("Z must be a one-dimensional torch tensor")
        nat = int(Z.shape[0])
        if force is None:
            force = torch.empty((nat, 3), dtype=torch.float64, device=getattr(pos, "device", None))
        _check_device_inputs(torch, nat, Z, pos, force)
        E, L = self.evaluate_device(nat, Z.data_ptr(), pos.data_ptr(), force.data_ptr(), periodic, lat, stream)
        return E, force, L

    # --- stage outputs of the last evaluation ------------------------------------
    def neighbor_entries(self) -> np.ndarray:
        n = ctypes.c_int64()
        self._check(self.lib.hdnnp_neighbor_count(self.ctx, ctypes.byref(n)))
        ent = np.zeros((max(n.value, 1), 5), dtype=np.int32)
        self._check(self.lib.hdnnp_get_neighbors(self.ctx, n.value, _p(ent, _ip)))
        return ent[:n.value]

    def symfunctions(self, nat: int):
        G = np.zeros((nat, self.maxsf))
        D = np.zeros((nat, self.maxsf))
        Ei = np.zeros(nat)
        self._check(self.lib.hdnnp_get_symfunctions(self.ctx, self.maxsf, _p(G, _dp), _p(D, _dp), _p(Ei, _dp)))
        return G, D, Ei

    def timings(self):
        """stage times (ms) of the last evaluation; zeros unless it ran with
        stage timings on (set_stage_timings)"""
        t = np.zeros(6)
        self._check(self.lib.hdnnp_last_timings(self.ctx, _p(t, _dp)))
        return t

    def set_stage_timings(self, on: bool):
        """hdnnp_set_stage_timings: HIP events between the stages (off by
        default; they cost a small evaluation ~10-20 us)"""
        self._check(self.lib.hdnnp_set_stage_timings(self.ctx, 1 if on else 0))


def parseRuNNerInput(path: str, precision="fp64") -> Hdnnp:
    return Hdnnp(path, precision)


class NeighborList:
    """type neighborList (neighborLists.f90:10-16): nNeis entries, slot 0 the
    atom itself; neis[n] the position of the image, ineis[n] its atom
    (0-based here, 1-based in the Fortran type)."""

    def __init__(self, neis, ineis, stage):
        self.nNeis = int(ineis.shape[0])
        self.neis = neis
        self.ineis = ineis
        self.stage = stage


class SymFunctionContainer:
    """type symFunctionContainer (symmetryFunctions.f90:24-33) of one atom:
    sf[s] scaled values, dsf[s, n] = dG_s/dR of list slot n (3 components)."""

    def __init__(self, sf, dsf, ctx, stage):
        self.nsf = int(sf.shape[0])
        self.sf = sf
        self.dsf = dsf
        self.ctx = ctx
        self.stage = stage


def _c_structure(ats):
    Z = np.ascontiguousarray(ats.Z, dtype=np.int32)
    pos = np.ascontiguousarray(ats.pos, dtype=np.float64).reshape(-1, 3)
    lat = np.ascontiguousarray(ats.lat, dtype=np.float64).reshape(3, 3)
    return Z, pos, int(bool(ats.periodic)), lat


def hdnnpCalcNeighborLists(handle: Hdnnp, ats):
    """hdnnpCalcNeighborLists (hdnnps.f90:373-388): the lists of every atom, in
    the reference's order and layout, built on the device."""
    Z, pos, per, lat = _c_structure(ats)
    nat = Z.shape[0]
    stage = ctypes.c_int64()
    nn = np.zeros(max(nat, 1), np.int32)
    handle._check(handle.lib.hdnnp_stage_neighbor_lists(handle.ctx, nat, _p(Z, _ip), _p(pos, _dp), per,
                                                        _p(lat, _dp), ctypes.byref(stage), _p(nn, _ip)))
    tot = int(nn[:nat].sum())
    ineis = np.zeros(max(tot, 1), np.int32)
    neis = np.zeros((max(tot, 1), 3))
    handle._check(handle.lib.hdnnp_stage_get_neighbor_lists(handle.ctx, stage.value, _p(ineis, _ip), _p(neis, _dp)))
    off = np.concatenate([[0], np.cumsum(nn[:nat])])
    return [NeighborList(neis[off[i]:off[i + 1]], ineis[off[i]:off[i + 1]] - 1, stage.value) for i in range(nat)]


def hdnnpCalcSymfunctions(handle: Hdnnp, ats, neiLists, calcSfsTogether: bool = True):
    """hdnnpCalcSymfunctions (hdnnps.f90:390-485): scaled values and dG/dR of
    every atom's symmetry functions. calcSfsTogether selects the fused pass
    (calcAllSymmetryFunctions) or the per-function kernel (symmetryFunction2
    / 3) for the values; the gradients come from the per-function kernel."""
    Z, pos, per, lat = _c_structure(ats)
    nat = Z.shape[0]
    rank = {z: e for e, z in enumerate(handle.elements)}
    nsf = np.array([handle.info.nsf[rank[int(z)]] for z in Z], dtype=np.int64)
    nn = np.array([nl.nNeis for nl in neiLists], dtype=np.int64)
    ld = max(handle.maxsf, 1)
    G = np.zeros((max(nat, 1), ld))
    off = np.concatenate([[0], np.cumsum(3 * nsf * nn)]).astype(np.int64)
    dsf = np.zeros(max(int(off[-1]), 1))
    stage = ctypes.c_int64(neiLists[0].stage if nat else 0)
    handle._check(handle.lib.hdnnp_stage_symfunctions(handle.ctx, ctypes.byref(stage), nat, _p(Z, _ip), _p(pos, _dp),
                                                      per, _p(lat, _dp), int(bool(calcSfsTogether)), ld,
                                                      _p(G, _dp), _p(dsf, _dp)))
    return [SymFunctionContainer(G[i, :nsf[i]].copy(), dsf[off[i]:off[i + 1]].reshape(nsf[i], nn[i], 3),
                                 handle.ctx, stage.value) for i in range(nat)]


def hdnnpCalcNNs(handle: Hdnnp, ats, neiLists, sfs, want_dEdLat: bool = True):
    """hdnnpCalcNNs (hdnnps.f90:487-562): this handle's networks on the
    symmetry functions sfs (computed by any handle of the same precision,
    elements and SF counts, as globalHdnnps.f90:53-55 does); forces and
    dEdLat from the gradients of the symmetry functions of the handle that
    computed sfs. Returns (E, F[nat,3], dEdLat|None)."""
    Z, pos, per, lat = _c_structure(ats)
    nat = Z.shape[0]
    ld = max(max((c.nsf for c in sfs), default=1), 1)
    G = np.zeros((max(nat, 1), ld))
    for i, c in enumerate(sfs):
        G[i, :c.nsf] = c.sf
    stage = ctypes.c_int64(sfs[0].stage if nat else 0)
    sfctx = sfs[0].ctx if nat else None
    E = ctypes.c_double()
    F = np.zeros((max(nat, 1), 3))
    L = np.zeros((3, 3)) if want_dEdLat else None
    handle._check(handle.lib.hdnnp_stage_nns(handle.ctx, sfctx, ctypes.byref(stage), nat, _p(Z, _ip), _p(pos, _dp),
                                             per, _p(lat, _dp), ld, _p(G, _dp), ctypes.byref(E), _p(F, _dp),
                                             _p(L, _dp)))
    return E.value, F[:nat], L


def hdnnpEnergyAndForces(handle: Hdnnp, ats, want_dEdLat: bool = True):
    return handle.energy_and_forces(ats, want_dEdLat)


# constants.f90 of the reference: mass unit and the standard atomic weights
# (amu) of the elements used here
MASS_UNIT_TO_ELECTRON_MASS = 1822.888513874368
FEMTOSECONDS_TO_ATOMIC = 41.34137333519408
_ATOMIC_MASS = {1: 1.008, 2: 4.002, 3: 6.94, 4: 9.012, 5: 10.81, 6: 12.011, 7: 14.007, 8: 15.999, 9: 18.998,
                10: 20.1797, 11: 22.989, 12: 24.305, 13: 26.981, 14: 28.085, 15: 30.973, 16: 32.06, 17: 35.45,
                18: 39.948, 53: 126.904, 82: 207.2}


def atomic_masses(Z) -> np.ndarray:
    """getAtomicMass(el) * massUnit2ElectronMass (electron masses)."""
    return np.array([_ATOMIC_MASS[int(z)] for z in Z]) * MASS_UNIT_TO_ELECTRON_MASS


class MdSystems:
    """hdnnp_md_*: velocity Verlet (hdnnpExample.f90:58-72) of several systems
    on the device with the forces of `handle`; momenta[k] (nat_k, 3) per system;
    skin > 0 reuses neighbour lists within rc + skin (Verlet skin)."""

    def __init__(self, handle: Hdnnp, structures, momenta, dt: float, skin: float = 0.0, masses=None):
        self.h = handle
        self.lib = handle.lib
        nat = np.array([s.nat for s in structures], dtype=np.int32)
        Z = np.ascontiguousarray(np.concatenate([np.asarray(s.Z, np.int32) for s in structures]))
        pos = np.ascontiguousarray(np.concatenate([np.asarray(s.pos, np.float64).reshape(-1, 3)
                                                   for s in structures]))
        mom = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in momenta]))
        mass = np.ascontiguousarray(atomic_masses(Z) if masses is None else np.asarray(masses, np.float64))
        per = np.array([int(bool(s.periodic)) for s in structures], dtype=np.int32)
        lat = np.ascontiguousarray(np.stack([np.asarray(s.lat, np.float64).reshape(3, 3) for s in structures]))
        self.nsys = len(structures)
        self.nat = int(nat.sum())
        self._off = np.concatenate([[0], np.cumsum(nat)])
        self.md = ctypes.c_void_p()
        handle._check(self.lib.hdnnp_md_create(handle.ctx, self.nsys, _p(nat, _ip), _p(Z, _ip), _p(pos, _dp),
                                               _p(mom, _dp), _p(mass, _dp), _p(per, _ip), _p(lat, _dp),
                                               float(dt), float(skin), ctypes.byref(self.md)))

    def step(self, n: int = 1):
        """n steps; returns (epot[nsys], ekin[nsys]) of the last one."""
        ep = np.zeros(self.nsys)
        ek = np.zeros(self.nsys)
        self.h._check(self.lib.hdnnp_md_step(self.md, n, _p(ep, _dp), _p(ek, _dp)))
        return ep, ek

    def get(self):
        pos = np.zeros((self.nat, 3))
        mom = np.zeros((self.nat, 3))
        f = np.zeros((self.nat, 3))
        self.h._check(self.lib.hdnnp_md_get(self.md, _p(pos, _dp), _p(mom, _dp), _p(f, _dp)))
        return pos, mom, f

    @property
    def stats(self):
        st, nb = ctypes.c_int64(), ctypes.c_int64()
        self.h._check(self.lib.hdnnp_md_stats(self.md, ctypes.byref(st), ctypes.byref(nb)))
        return st.value, nb.value

    @property
    def counters(self) -> dict:
        """steps, list builds, chunks cut back, graph captures, graph replays"""
        out = (ctypes.c_int64 * 5)()
        self.h._check(self.lib.hdnnp_md_counters(self.md, out))
        return dict(zip(("steps", "builds", "rollbacks", "graph_captures", "graph_replays"), list(out)))

    def close(self):
        if getattr(self, "md", None) and self.md.value:
            self.lib.hdnnp_md_destroy(self.md)
            self.md = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """hdnnp_server_*: a request server on the device. Evaluations are queued
    and evaluated in continuous batches by `lanes` worker threads (each with
    its own stream), MD sessions keep their state and neighbour lists on the
    device, and admission is by the device-memory budget (bytes; 0: 80% of
    the free memory). Results equal those of Hdnnp evaluations bit for bit."""

    def __init__(self, param_dir, precision="mixed", lanes: int = 4, hbm_budget: int = 0,
                 max_batch_atoms: int = 0):
        self.lib = load_library()
        dirs = list(param_dir) if isinstance(param_dir, (list, tuple)) else [param_dir]
        arr = (ctypes.c_char_p * len(dirs))(*[os.fsencode(d) for d in dirs])
        self.sv = ctypes.c_void_p()
        self._check(self.lib.hdnnp_server_create(len(dirs), arr, _PREC[precision], int(lanes), int(hbm_budget),
                                                 int(max_batch_atoms), ctypes.byref(self.sv)))
        self._nat = {}

    def _check(self, rc):
        if rc:
            raise HdnnpError(rc, self.lib.hdnnp_server_last_error().decode())

    def submit(self, s) -> int:
        """Queue the evaluation of structure s (.Z, .pos, .lat, .periodic); its ticket."""
        Z = np.ascontiguousarray(s.Z, dtype=np.int32)
        pos = np.ascontiguousarray(s.pos, dtype=np.float64).reshape(-1, 3)
        lat = np.ascontiguousarray(s.lat, dtype=np.float64).reshape(3, 3)
        t = ctypes.c_int64()
        self._check(self.lib.hdnnp_server_submit(self.sv, len(Z), _p(Z, _ip), _p(pos, _dp), int(bool(s.periodic)),
                                                 _p(lat, _dp), ctypes.byref(t)))
        self._nat[t.value] = len(Z)
        return t.value

    def result(self, ticket: int, timeout: float = 0.0):
        """(energy, forces (nat, 3), dEdLat (3, 3), energy std, latency ms) of a
        ticket; waits for it (timeout in seconds, 0: no limit)."""
        nat = self._nat.get(ticket)
        f = np.zeros((nat or 0, 3))
        d = np.zeros((3, 3))
        e, sd, lat = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        rc = self.lib.hdnnp_server_wait(self.sv, ticket, float(timeout), ctypes.byref(e),
                                        _p(f, _dp) if nat else None, _p(d, _dp), ctypes.byref(sd),
                                        ctypes.byref(lat))
        if rc != HDNNP_ERR_TIMEOUT:
            self._nat.pop(ticket, None)
        self._check(rc)
        return e.value, f, d, sd.value, lat.value

    def cancel(self, ticket: int) -> bool:
        """Cancel a queued evaluation or MD-step request: True when it had not
        started (waiting for it then raises HdnnpError with code
        HDNNP_ERR_CANCELLED), False when it runs or is done."""
        c = ctypes.c_int32()
        self._check(self.lib.hdnnp_server_cancel(self.sv, int(ticket), ctypes.byref(c)))
        return bool(c.value)

    def md_open(self, s, momenta, dt: float, skin: float = 1.0, masses=None) -> int:
        """Open an MD session of structure s with momenta (nat, 3); its id."""
        Z = np.ascontiguousarray(s.Z, dtype=np.int32)
        pos = np.ascontiguousarray(s.pos, dtype=np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(momenta, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(atomic_masses(Z) if masses is None else masses, dtype=np.float64)
        lat = np.ascontiguousarray(s.lat, dtype=np.float64).reshape(3, 3)
        sid = ctypes.c_int64()
        self._check(self.lib.hdnnp_server_md_open(self.sv, len(Z), _p(Z, _ip), _p(pos, _dp), _p(mom, _dp),
                                                  _p(mass, _dp), int(bool(s.periodic)), _p(lat, _dp), float(dt),
                                                  float(skin), ctypes.byref(sid)))
        self._nat[("s", sid.value)] = len(Z)
        return sid.value

    def md_step(self, session: int, nsteps: int) -> int:
        t = ctypes.c_int64()
        self._check(self.lib.hdnnp_server_md_step(self.sv, session, int(nsteps), ctypes.byref(t)))
        return t.value

    def md_result(self, ticket: int, timeout: float = 0.0):
        """(epot, ekin, latency ms) after an MD request."""
        ep, ek, lat = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.hdnnp_server_md_wait(self.sv, ticket, float(timeout), ctypes.byref(ep),
                                                  ctypes.byref(ek), ctypes.byref(lat)))
        return ep.value, ek.value, lat.value

    def md_get(self, session: int):
        nat = self._nat[("s", session)]
        pos, mom, f = np.zeros((nat, 3)), np.zeros((nat, 3)), np.zeros((nat, 3))
        self._check(self.lib.hdnnp_server_md_get(self.sv, session, _p(pos, _dp), _p(mom, _dp), _p(f, _dp)))
        return pos, mom, f

    def md_close(self, session: int):
        self._check(self.lib.hdnnp_server_md_close(self.sv, session))
        self._nat.pop(("s", session), None)

    def stats(self, reset: bool = False) -> dict:
        out = np.zeros(len(SERVER_STATS))
        self._check(self.lib.hdnnp_server_stats(self.sv, int(reset), _p(out, _dp)))
        return dict(zip(SERVER_STATS, out.tolist()))

    def close(self):
        if getattr(self, "sv", None) and self.sv.value:
            self.lib.hdnnp_server_destroy(self.sv)
            self.sv = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def initGlobalHdnnpsFromInputFile(path: str = "nnps.txt", precision="fp64") -> Hdnnp:
    """globalHdnnps.f90:13-29: the committee listed in `path` (first line the
    number of potentials, then one parameter directory per line)."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        dirs = [f.readline().strip() for _ in range(n)]
    return Hdnnp(dirs, precision)


def globalHdnnpEnergyAndForces(handle: Hdnnp, ats):
    """globalHdnnps.f90:31-70: committee mean E, F, dEdLat and the standard
    deviation of the member energies."""
    E, F, L = handle.energy_and_forces(ats)
    sd, _ = handle.committee(1)
    return E, F, L, float(sd[0])


def dEdLat2Stress(lat, dEdLat) -> np.ndarray:
    lib = load_library()
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    d = np.ascontiguousarray(dEdLat, dtype=np.float64)
    out = np.zeros((3, 3))
    rc = lib.hdnnp_dEdLat_to_stress(_p(lat, _dp), _p(d, _dp), _p(out, _dp))
    if rc:
        raise HdnnpError(rc, "dEdLat2Stress failed")
    return out
