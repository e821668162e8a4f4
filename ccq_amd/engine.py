"""ctypes binding of the HIP amplitude engine (include/ccq_hip.h).

This mirrors ccq's plug-in operator interface for the CC iteration:
``update`` is update_ptr (src/proc_pointers.f90:8-32 -> update_clusters_t4),
``energy`` is calculate_energy_ptr (calculate_unsorted_energy), and the DIIS
calls are write_vecs / residuum / calc_diis (src/diis.F90).  Errors raise
``EngineError`` carrying the library's message, like ccq's stop_all.

The library must have been built (``python ccq_amd/build.py`` or
``__graft_entry__.build()``); there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccq_hip.so")

LEVELS = {"CCSD": 2, "CCSDT": 3, "CCSDTQ": 4, "CCSDt": 5, "CC(t;3)": 6}


class EngineError(RuntimeError):
    pass


class Dims(ctypes.Structure):
    _fields_ = [("froz", ctypes.c_int64), ("occ_a", ctypes.c_int64),
                ("occ_b", ctypes.c_int64), ("orbs", ctypes.c_int64)]


class Opts(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int64), ("rhf", ctypes.c_int64), ("shift", ctypes.c_double),
                ("diis_space", ctypes.c_int64), ("device", ctypes.c_int64),
                ("act_occ", ctypes.c_int64), ("act_unocc", ctypes.c_int64)]


_lib = None

EXPORTS = {
    "ccq_hip_create": (ctypes.c_int, [ctypes.POINTER(Dims), ctypes.POINTER(Opts), ctypes.POINTER(ctypes.c_void_p)]),
    "ccq_hip_set_ints": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5),
    "ccq_hip_t_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "ccq_hip_set_t": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_get_t": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_t4_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "ccq_hip_set_t4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_get_t4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_vec_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "ccq_hip_write_vecs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_char_p]),
    "ccq_hip_read_vec": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "ccq_hip_write_t4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "ccq_hip_read_t4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "ccq_hip_update": (ctypes.c_int, [ctypes.c_void_p]),
    "ccq_hip_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_diis_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_residuum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_diis_extrapolate": (ctypes.c_int, [ctypes.c_void_p]),
    "ccq_hip_hbar": (ctypes.c_int, [ctypes.c_void_p]),
    "ccq_hip_left_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "ccq_hip_left_update": (ctypes.c_int, [ctypes.c_void_p]),
    "ccq_hip_left_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_left_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_l_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "ccq_hip_set_l": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_get_l": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_cct3": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_guess": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_guess_data_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "ccq_hip_guess_data": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 3 + [ctypes.c_int64]),
    "ccq_hip_mlp_nparams": (ctypes.c_int64, []),
    "ccq_hip_mlp_train": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_double, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_mlp_grad": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_mlp_grad_noalloc": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_mlp_last_error": (ctypes.c_char_p, []),
    "ccq_hip_onebody_orbitals": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "ccq_hip_read_onebody": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p]),
    "ccq_hip_read_twobody": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_fcidump_header": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(ctypes.c_int64)]),
    "ccq_hip_read_fcidump": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_read_sym": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]),
    "ccq_hip_read_last_error": (ctypes.c_char_p, []),
    "ccq_hip_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_profile_report": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "ccq_hip_debug_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_debug_sweep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_mfma_ceiling": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]),
    "ccq_hip_set_irreps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ccq_hip_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "ccq_hip_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ccq_hip_destroy": (ctypes.c_int, [ctypes.c_void_p]),
}


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"{path} is missing: build it with `python ccq_amd/build.py`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Engine:
    """One resident CC problem on one GPU."""

    def __init__(self, froz: int, occ_a: int, occ_b: int, orbs: int, calc_type: str = "CCSD",
                 rhf: bool = False, shift: float = 0.0, diis_space: int = 5, device: int = 0,
                 act_occ: int = -1, act_unocc: int = -1):
        self.lib = load_library()
        self.dims = Dims(froz, occ_a, occ_b, orbs)
        self.opts = Opts(LEVELS[calc_type], int(rhf), float(shift), int(diis_space), int(device),
                         int(act_occ), int(act_unocc))
        h = ctypes.c_void_p()
        rc = self.lib.ccq_hip_create(ctypes.byref(self.dims), ctypes.byref(self.opts), ctypes.byref(h))
        if rc != 0 or not h.value:
            raise EngineError("ccq_hip_create failed (see stderr)")
        self.h = h
        self.calc_type = calc_type
        self.t_size = int(self.lib.ccq_hip_t_size(self.h))
        self.t4_size = int(self.lib.ccq_hip_t4_size(self.h))
        self.l_size = int(self.lib.ccq_hip_l_size(self.h))

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.ccq_hip_last_error(self.h)
            raise EngineError(f"{what}: {msg.decode() if msg else 'error'}")

    def set_ints(self, f_a, f_b, v_aa, v_ab, v_bb):
        """Arrays are the logical ccq arrays f(p,q), v(p,q,r,s) (any numpy order)."""
        bufs = [np.asfortranarray(x, dtype=np.float64) for x in (f_a, f_b, v_aa, v_ab, v_bb)]
        if v_bb is v_aa:
            bufs[4] = bufs[2]
        self._check(self.lib.ccq_hip_set_ints(self.h, *[_ptr(b) for b in bufs]), "set_ints")

    def set_t(self, t_vec: np.ndarray):
        t = np.ascontiguousarray(t_vec, dtype=np.float64)
        self._check(self.lib.ccq_hip_set_t(self.h, _ptr(t), t.size), "set_t")

    def get_t(self) -> np.ndarray:
        out = np.empty(self.t_size)
        self._check(self.lib.ccq_hip_get_t(self.h, _ptr(out), out.size), "get_t")
        return out

    def set_t4(self, t4: np.ndarray):
        t = np.ascontiguousarray(t4, dtype=np.float64)
        self._check(self.lib.ccq_hip_set_t4(self.h, _ptr(t), t.size), "set_t4")

    def get_t4(self) -> np.ndarray:
        out = np.empty(self.t4_size)
        self._check(self.lib.ccq_hip_get_t4(self.h, _ptr(out), out.size), "get_t4")
        return out

    def write_vecs(self, it: int, diis_space: int, directory: str = "."):
        """ccq's write_vecs: t_vec.bin and cc_iter_mat-<slot>.bin in ``directory``."""
        self._check(self.lib.ccq_hip_write_vecs(self.h, it, diis_space, directory.encode()),
                    "write_vecs")

    def read_vec(self, path: str):
        """Restart: the DIIS part of t_vec from a t_vec.bin file."""
        self._check(self.lib.ccq_hip_read_vec(self.h, path.encode()), "read_vec")

    def write_t4(self, bin_file: str):
        self._check(self.lib.ccq_hip_write_t4(self.h, bin_file.encode()), "write_t4")

    def read_t4(self, bin_file: str):
        self._check(self.lib.ccq_hip_read_t4(self.h, bin_file.encode()), "read_t4")

    def update(self):
        self._check(self.lib.ccq_hip_update(self.h), "update")

    def energy(self) -> float:
        e = ctypes.c_double()
        self._check(self.lib.ccq_hip_energy(self.h, ctypes.byref(e)), "energy")
        return e.value

    def diis_record(self, it: int):
        self._check(self.lib.ccq_hip_diis_record(self.h, it), "diis_record")

    def residuum(self, it: int) -> float:
        r = ctypes.c_double()
        self._check(self.lib.ccq_hip_residuum(self.h, it, ctypes.byref(r)), "residuum")
        return r.value

    def diis_extrapolate(self):
        self._check(self.lib.ccq_hip_diis_extrapolate(self.h), "diis_extrapolate")

    def guess(self, mode: str, params=None):
        """Starting amplitudes: "zero", "mp2" or "learned" (params: packed float32)."""
        m = {"zero": 0, "mp2": 1, "learned": 2}[mode]
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.ccq_hip_guess(self.h, m, None if p is None else _ptr(p),
                                           0 if p is None else p.size), "guess")

    def guess_data(self):
        """(X[n, 5], M[n], Y[n]) of every T2 tuple at the current amplitudes."""
        n = int(self.lib.ccq_hip_guess_data_size(self.h))
        X = np.empty((n, 5), dtype=np.float32)
        M = np.empty(n, dtype=np.float32)
        Y = np.empty(n, dtype=np.float32)
        self._check(self.lib.ccq_hip_guess_data(self.h, _ptr(X), _ptr(M), _ptr(Y), n), "guess_data")
        return X, M, Y

    def set_precision(self, mode: int):
        """0: FP64 MFMA contractions; 1: FP32 MFMA for the tiled contractions;
        2: as 1, with every tensor of the sweep stored in FP32 as well (the
        updated amplitudes are FP64 in every mode)."""
        self._check(self.lib.ccq_hip_set_precision(self.h, int(mode)), "set_precision")

    # CC(t;3) tail (calc_type "CC(t;3)")

    def hbar(self):
        self._check(self.lib.ccq_hip_hbar(self.h), "hbar")

    def left_init(self, e_cor: float):
        self._check(self.lib.ccq_hip_left_init(self.h, float(e_cor)), "left_init")

    def left_update(self):
        self._check(self.lib.ccq_hip_left_update(self.h), "left_update")

    def left_energy(self) -> float:
        e = ctypes.c_double()
        self._check(self.lib.ccq_hip_left_energy(self.h, ctypes.byref(e)), "left_energy")
        return e.value

    def left_mode(self, on: bool):
        self._check(self.lib.ccq_hip_left_mode(self.h, int(on)), "left_mode")

    def set_l(self, l_vec: np.ndarray):
        x = np.ascontiguousarray(l_vec, dtype=np.float64)
        self._check(self.lib.ccq_hip_set_l(self.h, _ptr(x), x.size), "set_l")

    def get_l(self) -> np.ndarray:
        out = np.empty(self.l_size)
        self._check(self.lib.ccq_hip_get_l(self.h, _ptr(out), out.size), "get_l")
        return out

    def cct3(self):
        """crcc23's (e23a, e23b, e23c, e23d)."""
        e = (ctypes.c_double * 4)()
        self._check(self.lib.ccq_hip_cct3(self.h, e), "cct3")
        return tuple(e)

    def profile(self, enable: bool):
        self._check(self.lib.ccq_hip_profile(self.h, int(enable)), "profile")

    def set_irreps(self, irreps):
        """Orbital irreps, 1-based as ccq's sym_file (src/symmetry.f90:84-114),
        one per orbital; before set_ints.  Products then run in irrep blocks."""
        ir = np.ascontiguousarray(irreps, dtype=np.int64)
        self._check(self.lib.ccq_hip_set_irreps(self.h, _ptr(ir), ir.size), "set_irreps")

    def profile_report(self, sections: bool = False) -> dict:
        """Per launch class {calls, ms, flops, bytes} of the profiled launches;
        sections=True returns the per-section totals ("section:<name>") instead."""
        buf = ctypes.create_string_buffer(1 << 20)
        self._check(self.lib.ccq_hip_profile_report(self.h, buf, len(buf)), "profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, calls, ms, flops, nbytes = line.split()
            if name.startswith("section:") != sections:
                continue
            out[name] = dict(calls=int(calls), ms=float(ms), flops=float(flops), bytes=float(nbytes))
        return out

    def debug_read(self, tensor_id: str, shape) -> np.ndarray:
        out = np.empty(shape)
        self._check(self.lib.ccq_hip_debug_read(self.h, tensor_id.encode(), _ptr(out), out.size),
                    "debug_read")
        return out

    def debug_sweep(self, nsections: int):
        self._check(self.lib.ccq_hip_debug_sweep(self.h, nsections), "debug_sweep")

    def sync(self):
        self._check(self.lib.ccq_hip_sync(self.h), "sync")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ccq_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# t_vec block layout (ccq's cc%pos, src/utils/cc_utils.f90:479-559)
TVEC_BLOCKS = ['t1a', 't1b', 't2aa', 't2ab', 't2bb', 't3aaa', 't3aab', 't3abb', 't3bbb']


def block_shapes(oa, ob, va, vb):
    return {
        't1a': (oa, va), 't1b': (ob, vb),
        't2aa': (oa, oa, va, va), 't2ab': (oa, ob, va, vb), 't2bb': (ob, ob, vb, vb),
        't3aaa': (oa, oa, oa, va, va, va), 't3aab': (oa, oa, ob, va, va, vb),
        't3abb': (oa, ob, ob, va, vb, vb), 't3bbb': (ob, ob, ob, vb, vb, vb),
    }


T4_BLOCKS = ['t4aaaa', 't4aaab', 't4aabb', 't4abbb', 't4bbbb']


def t4_shapes(oa, ob, va, vb):
    return {
        't4aaaa': (oa, oa, oa, oa, va, va, va, va), 't4aaab': (oa, oa, oa, ob, va, va, va, vb),
        't4aabb': (oa, oa, ob, ob, va, va, vb, vb), 't4abbb': (oa, ob, ob, ob, va, vb, vb, vb),
        't4bbbb': (ob, ob, ob, ob, vb, vb, vb, vb),
    }


def split_t4(t4, oa, ob, va, vb) -> dict:
    out, off = {}, 0
    for name, shp in t4_shapes(oa, ob, va, vb).items():
        n = int(np.prod(shp))
        out[name] = t4[off:off + n].reshape(shp)
        off += n
    return out


def join_t4(blocks: dict, oa, ob, va, vb) -> np.ndarray:
    return np.concatenate([np.ascontiguousarray(blocks[n]).ravel()
                           for n in t4_shapes(oa, ob, va, vb)])


def split_tvec(t_vec, oa, ob, va, vb) -> dict:
    out = {}
    off = 0
    for name, shp in block_shapes(oa, ob, va, vb).items():
        n = int(np.prod(shp))
        out[name] = t_vec[off:off + n].reshape(shp)
        off += n
    return out


def join_tvec(blocks: dict, oa, ob, va, vb) -> np.ndarray:
    parts = []
    for name, shp in block_shapes(oa, ob, va, vb).items():
        b = blocks.get(name)
        parts.append(np.zeros(int(np.prod(shp))) if b is None else np.ascontiguousarray(b).ravel())
    return np.concatenate(parts)
