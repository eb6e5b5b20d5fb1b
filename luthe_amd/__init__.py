"""Python binding of libluthe.so, the MI355X TFHE bootstrapping engine.

The product is the C ABI declared in ``include/luthe.h``; this module binds it with
ctypes for tests, benchmarks and Python hosts.  Device buffers and events are taken
from the same HIP runtime libluthe links (``libamdhip64``), so one runtime serves the
whole process.  There is no CPU execution path: loading fails loudly when the
library has not been built, and ``Context`` fails when no gfx950 GPU is present.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libluthe.so")
_LIB = None
_HIP = None

P = 0xFFFFFFFF00000001
ERRORS = {0: "OK", -1: "ARG", -2: "PARAMS", -3: "HIP", -4: "NOMEM", -5: "DOP", -6: "RANGE",
          -7: "STATE", -8: "TIMEOUT", -9: "NODEV", -10: "BUSY"}
IOP_ADD, IOP_CMP_GT, IOP_MUL, IOP_ADD_RIPPLE = 0, 1, 2, 3
COHORT_OLDEST, COHORT_SHORTEST = 0, 1  # luthe_server_config.policy
IOP_NAMES = {IOP_ADD: "ADD", IOP_CMP_GT: "CMP_GT", IOP_MUL: "MUL", IOP_ADD_RIPPLE: "ADD_RIPPLE"}
IOP_LUTS = 11
BSK_COEF, BSK_NTT = 0, 1


class LutheError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"luthe error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("lwe_dim", C.c_uint32), ("glwe_k", C.c_uint32), ("poly_size", C.c_uint32),
                ("pbs_base_log", C.c_uint32), ("pbs_level", C.c_uint32), ("ks_base_log", C.c_uint32),
                ("ks_level", C.c_uint32), ("ks_mod_log", C.c_uint32), ("msg_bits", C.c_uint32),
                ("carry_bits", C.c_uint32), ("lwe_noise_std", C.c_double), ("glwe_noise_std", C.c_double),
                ("regfile_size", C.c_uint32), ("ct_slots", C.c_uint32), ("lut_slots", C.c_uint32)]

    @property
    def blen(self):
        return self.glwe_k * self.poly_size + 1

    @property
    def lut_len(self):
        return (self.glwe_k + 1) * self.poly_size

    def bsk_shape(self):
        k1 = self.glwe_k + 1
        return (self.lwe_dim, k1, self.pbs_level, k1, self.poly_size)

    def ksk_shape(self):
        return (self.glwe_k * self.poly_size, self.ks_level, self.lwe_dim + 1)


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("pbs", "pbs_batches", "pbs_max_batch", "alu", "mem", "dops",
                                          "iops", "tickets", "levels", "groups")]


class DopBatching(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("fill_pbs", "timeout_us", "max_groups")]


class IopInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("nblocks", "rows", "nodes", "npbs", "nalu", "depth", "nout",
                                          "a_row0", "b_row0")]


class IopNode(C.Structure):
    _fields_ = [("is_pbs", C.c_uint8), ("opcode", C.c_uint8), ("lut", C.c_uint8), ("log_lut_nb", C.c_uint8),
                ("dst", C.c_uint32), ("src0", C.c_uint32), ("src1", C.c_uint32), ("mul_factor", C.c_uint32)]


class HpuLayout(C.Structure):
    _fields_ = [("axi_bytes", C.c_uint32), ("pem_pc", C.c_uint32), ("regf_coef_nb", C.c_uint32),
                ("bsk_chunk", C.c_uint32), ("ntt_ncut", C.c_uint32), ("ntt_cut", C.c_uint32 * 8)]


class HpuKskLayout(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("axi_bytes", "ksk_pc", "ksk_cut_nb", "lbx", "lby", "lbz")]


def hpu_ksk_layout(axi_bytes=64, ksk_pc=1, ksk_cut_nb=1, lbx=3, lby=64, lbz=3):
    """KSK image geometry: AXI width, KSK_PC, KSK_CUT_NB and the key-switch block LBX x LBY x LBZ."""
    return HpuKskLayout(axi_bytes, ksk_pc, ksk_cut_nb, lbx, lby, lbz)


def hpu_layout(axi_bytes=64, pem_pc=2, regf_coef_nb=32, bsk_chunk=64, ntt_cut=(5, 6)):
    lay = HpuLayout()
    lay.axi_bytes, lay.pem_pc, lay.regf_coef_nb, lay.bsk_chunk = axi_bytes, pem_pc, regf_coef_nb, bsk_chunk
    lay.ntt_ncut = len(ntt_cut)
    for i, c in enumerate(ntt_cut):
        lay.ntt_cut[i] = c
    return lay


class TraceEvent(C.Structure):
    _fields_ = [("timestamp", C.c_uint32), ("insn", C.c_uint32)] + [
        (n, C.c_uint8) for n in ("cmd", "sync_id", "rd_lock", "wr_lock", "issue_lock", "vld", "rd_pdg", "pdg")] + [
        ("pe_reserved", C.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


TRC_NONE, TRC_RDUNLOCK, TRC_RETIRE, TRC_REFILL, TRC_ISSUE = 0, 1, 2, 3, 4


class Tuning(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("br_grid", "br_latency_max", "ntt_polys_per_wg", "cohort_max", "br_pair",
                                              "br_split_max", "br_split_wait_us", "ks_mfma_min")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class IopDopStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("dops", "ld", "st", "alu", "pbs", "heap_slots", "levels", "iop_opcode",
                                          "align_bits")]


class ServerConfig(C.Structure):
    _fields_ = [("arena_bytes", C.c_uint64), ("pinned_bytes", C.c_uint64), ("cohort_max", C.c_uint32),
                ("cohort_min", C.c_uint32), ("cohort_timeout_us", C.c_uint32), ("pipeline_depth", C.c_uint32),
                ("lut_base", C.c_uint32), ("max_inflight", C.c_uint32), ("queue_max", C.c_uint32),
                ("policy", C.c_uint32), ("age_limit_us", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = ([(n, C.c_uint64) for n in ("submitted", "admitted", "completed", "rejected", "steps", "cohorts",
                                           "cohorts_by_timeout", "pbs", "alu", "alu_launches")]
                + [("fill_hist", C.c_uint64 * 8)]
                + [(n, C.c_uint64) for n in ("arena_rows", "arena_rows_used", "arena_rows_peak")]
                + [(n, C.c_uint32) for n in ("cohort_max", "lut_base", "pending", "inflight")]
                + [(n, C.c_double) for n in ("latency_p50_us", "latency_p99_us", "latency_max_us")])

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "fill_hist"}
        d["fill_hist"] = list(self.fill_hist)
        return d


class RequestInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("op", "nblocks", "npbs", "depth", "cohorts")] + [
        ("status", C.c_int32), ("queue_us", C.c_double), ("latency_us", C.c_double)]


def lib():
    """Load libluthe.so (built by ``__graft_entry__.build()``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, sz, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_size_t, C.c_int
    P_ = C.POINTER(Params)
    sigs = {
        "luthe_params_default": (None, [P_]),
        "luthe_create": (i32, [P_, i32, C.POINTER(vp)]),
        "luthe_destroy": (i32, [vp]),
        "luthe_last_error": (C.c_char_p, [vp]),
        "luthe_get_params": (i32, [vp, P_]),
        "luthe_get_counters": (i32, [vp, C.POINTER(Counters)]),
        "luthe_device_bytes": (u64, [vp]),
        "luthe_upload_bsk": (i32, [vp, vp, i32]),
        "luthe_upload_ksk": (i32, [vp, vp]),
        "luthe_write_lut": (i32, [vp, u32, vp, u32]),
        "luthe_write_ct": (i32, [vp, u32, vp, u32]),
        "luthe_read_ct": (i32, [vp, u32, vp, u32]),
        "luthe_write_reg": (i32, [vp, u32, vp, u32]),
        "luthe_read_reg": (i32, [vp, u32, vp, u32]),
        "luthe_submit_dops": (i32, [vp, vp, sz, C.POINTER(u64)]),
        "luthe_poll": (i32, [vp, u64]),
        "luthe_wait": (i32, [vp, u64, C.c_int64]),
        "luthe_register_iop": (i32, [vp, C.c_uint8, C.c_uint8, vp, sz]),
        "luthe_submit_iop": (i32, [vp, vp, sz, C.POINTER(u64)]),
        "luthe_pbs_batch": (i32, [vp, vp, vp, vp, u32, sz, vp]),
        "luthe_pbs_batch_noalloc": (i32, [vp, vp, vp, vp, u32, sz, vp]),
        "luthe_pbs_batch_list": (i32, [vp, vp, u32, sz, vp]),
        "luthe_flush": (i32, [vp]),
        "luthe_set_dop_batching": (i32, [vp, C.POINTER(DopBatching)]),
        "luthe_get_dop_batching": (i32, [vp, C.POINTER(DopBatching)]),
        "luthe_keyswitch_batch": (i32, [vp, vp, vp, sz, vp]),
        "luthe_blind_rotate_batch": (i32, [vp, vp, vp, vp, u32, sz, vp]),
        "luthe_ntt_polymul_batch": (i32, [vp, vp, vp, vp, sz, vp]),
        "luthe_ntt_forward_batch": (i32, [vp, vp, vp, sz, vp]),
        "luthe_ntt_inverse_batch": (i32, [vp, vp, vp, sz, vp]),
        "luthe_sync": (i32, [vp]),
        "luthe_stream": (vp, [vp]),
        "luthe_set_br_grid": (i32, [vp, i32]),
        "luthe_get_br_grid": (i32, [vp]),
        "luthe_set_br_latency_max": (i32, [vp, i32]),
        "luthe_autotune": (i32, [vp, u32, C.c_char_p, C.POINTER(Tuning)]),
        "luthe_load_tuning": (i32, [vp, C.c_char_p]),
        "luthe_get_tuning": (i32, [vp, C.POINTER(Tuning)]),
        "luthe_set_tuning": (i32, [vp, C.POINTER(Tuning)]),
        "luthe_get_br_latency_max": (i32, [vp]),
        "luthe_set_br_split_local": (i32, [vp, i32]),
        "luthe_get_br_split_local": (i32, [vp]),
        "luthe_microbench_modmul": (i32, [vp, C.POINTER(C.c_double)]),
        "luthe_dop_schedule": (i32, [vp, sz, u32, u32, u32, vp]),
        "luthe_dop_plan": (i32, [vp, sz, u32, u32, u32, vp, vp, sz, C.POINTER(sz), C.POINTER(u32)]),
        "luthe_alu_batch": (i32, [vp, u32, vp, vp, vp, u32, u32, sz, vp]),
        "luthe_reserve": (i32, [vp, sz]),
        "luthe_capture_begin": (i32, [vp]),
        "luthe_capture_end": (i32, [vp, C.POINTER(vp)]),
        "luthe_graph_launch": (i32, [vp, vp]),
        "luthe_graph_destroy": (i32, [vp]),
        "luthe_iop_translate": (i32, [vp, sz, vp, sz, u32, vp]),
        "luthe_keygen_secret": (i32, [P_, u64, vp, vp]),
        "luthe_keygen_bsk": (i32, [P_, u64, vp, vp, vp, i32]),
        "luthe_keygen_ksk": (i32, [P_, u64, vp, vp, vp]),
        "luthe_encrypt_blwe": (i32, [P_, u64, vp, vp, sz, vp]),
        "luthe_decrypt_blwe": (i32, [P_, vp, vp, sz, vp]),
        "luthe_delta": (u64, [P_]),
        "luthe_decode": (u32, [P_, u64]),
        "luthe_lut_encode": (i32, [P_, vp, u32, vp]),
        "luthe_lut_encode_many": (i32, [P_, vp, u32, u32, vp]),
        "luthe_hpu_bsk_order": (i32, [vp, u32, vp]),
        "luthe_hpu_bsk_image_words": (sz, [P_]),
        "luthe_hpu_bsk_from_image": (i32, [P_, C.POINTER(HpuLayout), vp, vp]),
        "luthe_hpu_bsk_to_image": (i32, [P_, C.POINTER(HpuLayout), vp, vp]),
        "luthe_upload_bsk_hpu_image": (i32, [vp, C.POINTER(HpuLayout), vp]),
        "luthe_hpu_ct_slot_bytes": (sz, [P_, C.POINTER(HpuLayout)]),
        "luthe_hpu_ct_to_image": (i32, [P_, C.POINTER(HpuLayout), vp, u32, vp]),
        "luthe_hpu_ct_from_image": (i32, [P_, C.POINTER(HpuLayout), vp, u32, vp]),
        "luthe_hpu_lut_from_image": (i32, [P_, vp, u32, vp]),
        "luthe_write_lut_hpu_image": (i32, [vp, u32, vp, u32]),
        "luthe_hpu_ksk_image_bytes": (sz, [P_, C.POINTER(HpuKskLayout)]),
        "luthe_hpu_ksk_to_image": (i32, [P_, C.POINTER(HpuKskLayout), vp, C.POINTER(vp)]),
        "luthe_hpu_ksk_from_image": (i32, [P_, C.POINTER(HpuKskLayout), C.POINTER(vp), vp]),
        "luthe_upload_ksk_hpu_image": (i32, [vp, C.POINTER(HpuKskLayout), C.POINTER(vp)]),
        "luthe_trace_record_bits": (i32, [u32]),
        "luthe_trace_pack": (i32, [u32, vp, sz, vp]),
        "luthe_trace_unpack": (i32, [u32, vp, sz, vp]),
        "luthe_trace_enable": (i32, [vp, sz]),
        "luthe_trace_read": (i32, [vp, vp, sz, C.POINTER(sz)]),
        "luthe_iop_program": (i32, [u32, u32, C.POINTER(IopInfo), vp, sz, vp, sz]),
        "luthe_iop_lut": (i32, [u32, vp, C.POINTER(u32)]),
        "luthe_iop_program_dops": (i32, [u32, u32, u32, u32, vp, sz, C.POINTER(sz), C.POINTER(IopDopStats)]),
        "luthe_write_iop_luts": (i32, [vp, u32]),
        "luthe_register_iop_program": (i32, [vp, u32, u32, u32]),
        "luthe_iop_eval_plain": (i32, [u32, u32, vp, vp, vp]),
        "luthe_server_config_default": (None, [C.POINTER(ServerConfig)]),
        "luthe_server_create": (i32, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
        "luthe_server_destroy": (i32, [vp]),
        "luthe_server_last_error": (C.c_char_p, [vp]),
        "luthe_server_submit": (i32, [vp, u32, u32, vp, vp, vp, C.POINTER(u64)]),
        "luthe_server_poll": (i32, [vp, u64, C.POINTER(RequestInfo)]),
        "luthe_server_wait": (i32, [vp, u64, C.c_int64, C.POINTER(RequestInfo)]),
        "luthe_server_flush": (i32, [vp]),
        "luthe_server_get_stats": (i32, [vp, C.POINTER(ServerStats), i32]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    """Names of the C-ABI entry points this binding expects."""
    lib()
    return [n for n in dir(_LIB) if n.startswith("luthe_")]


def hip():
    """The HIP runtime libluthe.so is linked against (already loaded by it)."""
    global _HIP
    if _HIP is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        vp, sz = C.c_void_p, C.c_size_t
        for name, res, args in [
            ("hipMalloc", C.c_int, [C.POINTER(vp), sz]),
            ("hipFree", C.c_int, [vp]),
            ("hipMemcpy", C.c_int, [vp, vp, sz, C.c_int]),
            ("hipMemset", C.c_int, [vp, C.c_int, sz]),
            ("hipDeviceSynchronize", C.c_int, []),
            ("hipStreamSynchronize", C.c_int, [vp]),
            ("hipEventCreate", C.c_int, [C.POINTER(vp)]),
            ("hipEventDestroy", C.c_int, [vp]),
            ("hipEventRecord", C.c_int, [vp, vp]),
            ("hipEventSynchronize", C.c_int, [vp]),
            ("hipEventElapsedTime", C.c_int, [C.POINTER(C.c_float), vp, vp]),
            ("hipGetDeviceCount", C.c_int, [C.POINTER(C.c_int)]),
        ]:
            f = getattr(H, name)
            f.restype = res
            f.argtypes = args
        _HIP = H
    return _HIP


def _hipchk(rc, what):
    if rc != 0:
        raise LutheError(-3, f"{what} returned hipError {rc}")


def device_count():
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """A hipMalloc'd buffer viewed as a typed array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        _hipchk(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)), "hipMalloc")

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.shape, arr.dtype)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        _hipchk(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), self.nbytes, 1), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _hipchk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _hipchk(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def zero(self):
        _hipchk(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _vp(a):
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, C.c_void_p):
        return a
    if isinstance(a, int):
        return C.c_void_p(a)
    return a.ctypes.data_as(C.c_void_p)


# luthe_pbs_batch_desc: d_in, d_out, d_gid (device addresses), n, log_lut_nb
PBS_BATCH_DESC = np.dtype([("d_in", "<u8"), ("d_out", "<u8"), ("d_gid", "<u8"), ("n", "<u4"), ("log_lut_nb", "<u4")])


def batch_descs(batches):
    """Descriptor array for luthe_pbs_batch_list from (in_addr, out_addr, gid_addr, n, log_lut_nb) tuples."""
    d = np.zeros(len(batches), PBS_BATCH_DESC)
    for i, (a, o, g, n, lg) in enumerate(batches):
        d[i] = (a, o, g, n, lg)
    return d


def default_params(**over):
    p = Params()
    lib().luthe_params_default(C.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


class Context:
    """An engine context on one GPU (owns register file, memories and keys)."""

    def __init__(self, params=None, device=0):
        self.params = params if params is not None else default_params()
        self._h = C.c_void_p()
        rc = lib().luthe_create(C.byref(self.params), device, C.byref(self._h))
        if rc != 0:
            raise LutheError(rc, "luthe_create")

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().luthe_last_error(self._h)
            raise LutheError(rc, f"{what}: {msg.decode() if msg else ''}")

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return lib().luthe_stream(self._h)

    def close(self):
        if self._h:
            rc = lib().luthe_destroy(self._h)
            if rc == -7:  # a server is still attached: keep the context alive
                self._chk(rc, "luthe_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_bsk(self, bsk, layout=BSK_COEF):
        bsk = np.ascontiguousarray(bsk, np.uint64)
        assert bsk.size == int(np.prod(self.params.bsk_shape()))
        self._chk(lib().luthe_upload_bsk(self._h, _vp(bsk), layout), "upload_bsk")

    def upload_ksk(self, ksk):
        ksk = np.ascontiguousarray(ksk, np.uint32)
        assert ksk.size == int(np.prod(self.params.ksk_shape()))
        self._chk(lib().luthe_upload_ksk(self._h, _vp(ksk)), "upload_ksk")

    def write_lut(self, gid, glwe):
        glwe = np.ascontiguousarray(glwe, np.uint64).reshape(-1, self.params.lut_len)
        self._chk(lib().luthe_write_lut(self._h, gid, _vp(glwe), glwe.shape[0]), "write_lut")

    def write_ct(self, cid, blwe):
        blwe = np.ascontiguousarray(blwe, np.uint64).reshape(-1, self.params.blen)
        self._chk(lib().luthe_write_ct(self._h, cid, _vp(blwe), blwe.shape[0]), "write_ct")

    def read_ct(self, cid, count=1):
        out = np.empty((count, self.params.blen), np.uint64)
        self._chk(lib().luthe_read_ct(self._h, cid, _vp(out), count), "read_ct")
        return out

    def write_reg(self, rid, blwe):
        blwe = np.ascontiguousarray(blwe, np.uint64).reshape(-1, self.params.blen)
        self._chk(lib().luthe_write_reg(self._h, rid, _vp(blwe), blwe.shape[0]), "write_reg")

    def read_reg(self, rid, count=1):
        out = np.empty((count, self.params.blen), np.uint64)
        self._chk(lib().luthe_read_reg(self._h, rid, _vp(out), count), "read_reg")
        return out

    def submit_dops(self, words):
        w = np.ascontiguousarray(words, np.uint32)
        t = C.c_uint64(0)
        self._chk(lib().luthe_submit_dops(self._h, _vp(w), w.size, C.byref(t)), "submit_dops")
        return t.value

    def register_iop(self, opcode, max_align, words):
        w = np.ascontiguousarray(words, np.uint32)
        self._chk(lib().luthe_register_iop(self._h, opcode, max_align, _vp(w), w.size), "register_iop")

    def submit_iop(self, words):
        w = np.ascontiguousarray(words, np.uint32)
        t = C.c_uint64(0)
        self._chk(lib().luthe_submit_iop(self._h, _vp(w), w.size, C.byref(t)), "submit_iop")
        return t.value

    def poll(self, ticket):
        rc = lib().luthe_poll(self._h, ticket)
        if rc < 0:
            self._chk(rc, "poll")
        return rc == 1

    def wait(self, ticket, timeout_us=-1):
        self._chk(lib().luthe_wait(self._h, ticket, timeout_us), "wait")

    def flush(self):
        """Launch the open group of DOp/IOp submissions now."""
        self._chk(lib().luthe_flush(self._h), "flush")

    def set_dop_batching(self, fill_pbs=0, timeout_us=0, max_groups=1):
        b = DopBatching(fill_pbs, timeout_us, max_groups)
        self._chk(lib().luthe_set_dop_batching(self._h, C.byref(b)), "set_dop_batching")

    def dop_batching(self):
        b = DopBatching()
        self._chk(lib().luthe_get_dop_batching(self._h, C.byref(b)), "get_dop_batching")
        return {n: getattr(b, n) for n, _ in DopBatching._fields_}

    def sync(self):
        self._chk(lib().luthe_sync(self._h), "sync")

    def counters(self):
        c = Counters()
        self._chk(lib().luthe_get_counters(self._h, C.byref(c)), "counters")
        return {n: getattr(c, n) for n, _ in Counters._fields_}

    def device_bytes(self):
        return lib().luthe_device_bytes(self._h)

    def set_br_grid(self, wg):
        self._chk(lib().luthe_set_br_grid(self._h, wg), "set_br_grid")

    def br_grid(self):
        return lib().luthe_get_br_grid(self._h)

    def modmul_peak(self):
        """Measured Goldilocks modmul throughput at full occupancy, Gmodmul/s."""
        v = C.c_double(0)
        self._chk(lib().luthe_microbench_modmul(self._h, C.byref(v)), "microbench_modmul")
        return v.value

    # ---- batched device-memory entry points
    def autotune(self, batch, json_path=None):
        t = Tuning()
        path = json_path.encode() if json_path else None
        self._chk(lib().luthe_autotune(self._h, batch, path, C.byref(t)), "autotune")
        return t.as_dict()

    def load_tuning(self, json_path):
        self._chk(lib().luthe_load_tuning(self._h, json_path.encode()), "load_tuning")

    def tuning(self):
        t = Tuning()
        self._chk(lib().luthe_get_tuning(self._h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        self._chk(lib().luthe_get_tuning(self._h, C.byref(t)), "get_tuning")
        for k, v in kw.items():
            setattr(t, k, v)
        self._chk(lib().luthe_set_tuning(self._h, C.byref(t)), "set_tuning")

    def write_iop_luts(self, lut_base):
        self._chk(lib().luthe_write_iop_luts(self._h, lut_base), "write_iop_luts")

    def register_iop_program(self, op, nblocks, lut_base):
        self._chk(lib().luthe_register_iop_program(self._h, op, nblocks, lut_base), "register_iop_program")

    def trace_enable(self, capacity):
        self._chk(lib().luthe_trace_enable(self._h, capacity), "trace_enable")

    def trace_read(self, cap=1 << 16):
        buf = (TraceEvent * cap)()
        n = C.c_size_t()
        self._chk(lib().luthe_trace_read(self._h, C.cast(buf, C.c_void_p), cap, C.byref(n)), "trace_read")
        return [buf[i].as_dict() for i in range(n.value)]

    def upload_bsk_hpu_image(self, layout, image):
        img = np.ascontiguousarray(image, np.uint64)
        self._chk(lib().luthe_upload_bsk_hpu_image(self._h, C.byref(layout), _vp(img)), "upload_bsk_hpu_image")

    def upload_ksk_hpu_image(self, layout, images):
        imgs = [np.ascontiguousarray(i, np.uint8) for i in images]
        ptrs = (C.c_void_p * len(imgs))(*[i.ctypes.data for i in imgs])
        self._chk(lib().luthe_upload_ksk_hpu_image(self._h, C.byref(layout), ptrs), "upload_ksk_hpu_image")

    def write_lut_hpu_image(self, gid, bodies):
        b = np.ascontiguousarray(bodies, np.uint64).reshape(-1, self.params.poly_size)
        self._chk(lib().luthe_write_lut_hpu_image(self._h, gid, _vp(b), b.shape[0]), "write_lut_hpu_image")

    def alu_batch_dev(self, opcode, d_a, d_b, d_out, n, mul_factor=0, msg_cst=0, stream=None):
        self._chk(lib().luthe_alu_batch(self._h, opcode, _vp(d_a), _vp(d_b) if d_b is not None else None,
                                        _vp(d_out), mul_factor, msg_cst, n, stream), "alu_batch")

    def reserve(self, max_batch):
        self._chk(lib().luthe_reserve(self._h, max_batch), "reserve")

    def capture_begin(self):
        self._chk(lib().luthe_capture_begin(self._h), "capture_begin")

    def capture_end(self):
        g = C.c_void_p()
        self._chk(lib().luthe_capture_end(self._h, C.byref(g)), "capture_end")
        return Graph(self, g)

    def pbs_batch_dev(self, d_in, d_out, d_gid, n, log_lut_nb=0, stream=None):
        self._chk(lib().luthe_pbs_batch(self._h, _vp(d_in), _vp(d_out), _vp(d_gid), log_lut_nb, n, stream),
                  "pbs_batch")

    def pbs_batch_list_dev(self, d_descs, count, total, stream=None):
        """One KS + blind-rotation launch over `count` batches described by PBS_BATCH_DESC
        records in device memory (see batch_descs)."""
        self._chk(lib().luthe_pbs_batch_list(self._h, _vp(d_descs), count, total, stream), "pbs_batch_list")

    def pbs_batch(self, blwe, gids, log_lut_nb=0):
        """Host convenience: upload, run luthe_pbs_batch, download."""
        blwe = np.ascontiguousarray(blwe, np.uint64).reshape(-1, self.params.blen)
        n = blwe.shape[0]
        d_in = DeviceBuffer.from_numpy(blwe)
        d_gid = DeviceBuffer.from_numpy(np.ascontiguousarray(gids, np.uint32).reshape(n))
        d_out = DeviceBuffer((n << log_lut_nb, self.params.blen), np.uint64)
        self.pbs_batch_dev(d_in, d_out, d_gid, n, log_lut_nb)
        self.sync()
        return d_out.numpy()

    def keyswitch_batch(self, blwe):
        blwe = np.ascontiguousarray(blwe, np.uint64).reshape(-1, self.params.blen)
        n = blwe.shape[0]
        d_in = DeviceBuffer.from_numpy(blwe)
        d_out = DeviceBuffer((n, self.params.lwe_dim + 1), np.uint16)
        self._chk(lib().luthe_keyswitch_batch(self._h, d_in.ptr, d_out.ptr, n, None), "keyswitch_batch")
        self.sync()
        return d_out.numpy()

    def blind_rotate_batch(self, lwe2n, gids, log_lut_nb=0):
        lwe2n = np.ascontiguousarray(lwe2n, np.uint16).reshape(-1, self.params.lwe_dim + 1)
        n = lwe2n.shape[0]
        d_in = DeviceBuffer.from_numpy(lwe2n)
        d_gid = DeviceBuffer.from_numpy(np.ascontiguousarray(gids, np.uint32).reshape(n))
        d_out = DeviceBuffer((n << log_lut_nb, self.params.blen), np.uint64)
        self._chk(lib().luthe_blind_rotate_batch(self._h, d_in.ptr, d_out.ptr, d_gid.ptr, log_lut_nb, n, None),
                  "blind_rotate_batch")
        self.sync()
        return d_out.numpy()

    def ntt_polymul(self, a, bn):
        a = np.ascontiguousarray(a, np.uint64).reshape(-1, 2048)
        bn = np.ascontiguousarray(bn, np.uint64).reshape(-1, 2048)
        d_a, d_b = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(bn)
        d_o = DeviceBuffer(a.shape, np.uint64)
        self._chk(lib().luthe_ntt_polymul_batch(self._h, d_a.ptr, d_b.ptr, d_o.ptr, a.shape[0], None), "ntt_polymul")
        self.sync()
        return d_o.numpy()

    def ntt_forward(self, a):
        a = np.ascontiguousarray(a, np.uint64).reshape(-1, 2048)
        d_a, d_o = DeviceBuffer.from_numpy(a), DeviceBuffer(a.shape, np.uint64)
        self._chk(lib().luthe_ntt_forward_batch(self._h, d_a.ptr, d_o.ptr, a.shape[0], None), "ntt_forward")
        self.sync()
        return d_o.numpy()

    def ntt_inverse(self, a):
        a = np.ascontiguousarray(a, np.uint64).reshape(-1, 2048)
        d_a, d_o = DeviceBuffer.from_numpy(a), DeviceBuffer(a.shape, np.uint64)
        self._chk(lib().luthe_ntt_inverse_batch(self._h, d_a.ptr, d_o.ptr, a.shape[0], None), "ntt_inverse")
        self.sync()
        return d_o.numpy()


class Graph:
    """A captured sequence of batch calls, replayed on the context stream."""

    def __init__(self, ctx, handle):
        self.ctx, self._g = ctx, handle

    def launch(self):
        self.ctx._chk(lib().luthe_graph_launch(self.ctx.handle, self._g), "graph_launch")

    def destroy(self):
        if self._g:
            lib().luthe_graph_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---------------------------------------------------------------- client-side helpers
class Keys:
    """Seeded key material (secret keys, BSK in coefficient domain, KSK)."""

    def __init__(self, params, seed=1, threads=0, with_bsk=True):
        self.params = params
        L = lib()
        self.lwe_sk = np.zeros(params.lwe_dim, np.uint64)
        self.glwe_sk = np.zeros(params.glwe_k * params.poly_size, np.uint64)
        assert L.luthe_keygen_secret(C.byref(params), seed, _vp(self.lwe_sk), _vp(self.glwe_sk)) == 0
        self.ksk = np.zeros(params.ksk_shape(), np.uint32)
        assert L.luthe_keygen_ksk(C.byref(params), seed, _vp(self.lwe_sk), _vp(self.glwe_sk), _vp(self.ksk)) == 0
        self.bsk = None
        if with_bsk:
            self.bsk = np.zeros(params.bsk_shape(), np.uint64)
            assert L.luthe_keygen_bsk(C.byref(params), seed, _vp(self.lwe_sk), _vp(self.glwe_sk), _vp(self.bsk),
                                      threads) == 0

    def encrypt(self, plaintexts, seed=7):
        pt = np.ascontiguousarray(plaintexts, np.uint64).reshape(-1)
        out = np.empty((pt.size, self.params.blen), np.uint64)
        assert lib().luthe_encrypt_blwe(C.byref(self.params), seed, _vp(self.glwe_sk), _vp(pt), pt.size,
                                        _vp(out)) == 0
        return out

    def encrypt_msg(self, msgs, seed=7):
        d = lib().luthe_delta(C.byref(self.params))
        m = np.ascontiguousarray(msgs, np.uint64).reshape(-1)
        return self.encrypt((m * np.uint64(d)).astype(np.uint64), seed)

    def phase(self, blwe):
        blwe = np.ascontiguousarray(blwe, np.uint64).reshape(-1, self.params.blen)
        out = np.empty(blwe.shape[0], np.uint64)
        assert lib().luthe_decrypt_blwe(C.byref(self.params), _vp(self.glwe_sk), _vp(blwe), blwe.shape[0],
                                        _vp(out)) == 0
        return out

    def decrypt_msg(self, blwe):
        return np.array([lib().luthe_decode(C.byref(self.params), int(v)) for v in self.phase(blwe)], np.uint32)


def lut_encode(params, table):
    t = np.ascontiguousarray(table, np.uint8)
    out = np.empty(params.lut_len, np.uint64)
    rc = lib().luthe_lut_encode(C.byref(params), _vp(t), t.size, _vp(out))
    if rc != 0:
        raise LutheError(rc, "lut_encode")
    return out


def lut_encode_many(params, tables, log_lut_nb):
    t = np.ascontiguousarray(tables, np.uint8)
    out = np.empty(params.lut_len, np.uint64)
    rc = lib().luthe_lut_encode_many(C.byref(params), _vp(t), t.shape[-1], log_lut_nb, _vp(out))
    if rc != 0:
        raise LutheError(rc, "lut_encode_many")
    return out


def dop_schedule(words, regfile_size=64, ct_slots=4096, lut_slots=256):
    """Dependency levels of a concrete DOp stream (host only).  Returns (nlevels, levels)."""
    w = np.ascontiguousarray(words, np.uint32)
    lv = np.zeros(max(w.size, 1), np.int32)
    rc = lib().luthe_dop_schedule(_vp(w), w.size, regfile_size, ct_slots, lut_slots, _vp(lv))
    if rc < 0:
        raise LutheError(rc, "dop_schedule")
    return rc, lv[:w.size]


def dop_plan(words, regfile_size=64, ct_slots=4096, lut_slots=256):
    """The executor's renamed plan (host only): (nlevels, entries[n, 4] = level, dst, src0,
    src1 rows, copy_back[k, 2] = (register, row), rename_rows)."""
    w = np.ascontiguousarray(words, np.uint32)
    raw = np.zeros((max(w.size, 1), 4), np.uint32)
    ncb, rows = C.c_size_t(), C.c_uint32()
    cb = np.zeros(2 * regfile_size, np.uint32)
    rc = lib().luthe_dop_plan(_vp(w), w.size, regfile_size, ct_slots, lut_slots, _vp(raw), _vp(cb), regfile_size,
                              C.byref(ncb), C.byref(rows))
    if rc < 0:
        raise LutheError(rc, "dop_plan")
    ent = raw[:w.size].astype(np.int64)
    ent[:, 0] = raw[:w.size, 0].view(np.int32)
    return rc, ent, cb[:2 * ncb.value].reshape(-1, 2), rows.value


def iop_translate(iop_words, template, ct_slots=4096):
    """Patch a DOp template program for one IOp (host only)."""
    iw = np.ascontiguousarray(iop_words, np.uint32)
    tp = np.ascontiguousarray(template, np.uint32)
    out = np.zeros(max(tp.size, 1), np.uint32)
    rc = lib().luthe_iop_translate(_vp(iw), iw.size, _vp(tp), tp.size, ct_slots, _vp(out))
    if rc != 0:
        raise LutheError(rc, "iop_translate")
    return out[:tp.size]


def delta(params):
    return lib().luthe_delta(C.byref(params))


# ---------------------------------------------------------------- radix IOp programs
def iop_program(op, nblocks):
    """(info, nodes, out_rows) of a builder-written radix IOp program (host only)."""
    info = IopInfo()
    rc = lib().luthe_iop_program(op, nblocks, C.byref(info), None, 0, None, 0)
    if rc != 0:
        raise LutheError(rc, "iop_program")
    nodes = (IopNode * info.nodes)()
    outs = np.zeros(info.nout, np.uint32)
    rc = lib().luthe_iop_program(op, nblocks, C.byref(info), C.cast(nodes, C.c_void_p), info.nodes, _vp(outs),
                                 outs.size)
    if rc != 0:
        raise LutheError(rc, "iop_program")
    return info, nodes, outs


def iop_program_dops(op, nblocks, regs=64, lut_base=0):
    """(DOp template words, stats dict) of a radix IOp program (host only)."""
    n = C.c_size_t()
    st = IopDopStats()
    rc = lib().luthe_iop_program_dops(op, nblocks, regs, lut_base, None, 0, C.byref(n), C.byref(st))
    if rc != 0:
        raise LutheError(rc, "iop_program_dops")
    w = np.zeros(n.value, np.uint32)
    rc = lib().luthe_iop_program_dops(op, nblocks, regs, lut_base, _vp(w), w.size, C.byref(n), C.byref(st))
    if rc != 0:
        raise LutheError(rc, "iop_program_dops")
    return w, {k: getattr(st, k) for k, _ in st._fields_}


def iop_lut(lut):
    """(table of 16 entries, log_lut_nb) of a program LUT."""
    t = np.zeros(16, np.uint8)
    lognb = C.c_uint32()
    rc = lib().luthe_iop_lut(lut, _vp(t), C.byref(lognb))
    if rc != 0:
        raise LutheError(rc, "iop_lut")
    return t, lognb.value


def iop_eval_plain(op, a_blocks, b_blocks):
    a = np.ascontiguousarray(a_blocks, np.uint32)
    b = np.ascontiguousarray(b_blocks, np.uint32)
    out = np.zeros(1 if op == IOP_CMP_GT else a.size, np.uint32)
    rc = lib().luthe_iop_eval_plain(op, a.size, _vp(a), _vp(b), _vp(out))
    if rc != 0:
        raise LutheError(rc, "iop_eval_plain")
    return out


def to_blocks(x, nblocks):
    """Radix decomposition of an integer into 2-bit blocks, least significant first."""
    return np.array([(int(x) >> (2 * i)) & 3 for i in range(nblocks)], np.uint32)


def from_blocks(blocks):
    return sum(int(v) << (2 * i) for i, v in enumerate(blocks))


# ---------------------------------------------------------------- serving engine
class Server:
    """Continuous-batching server of radix IOps attached to a Context (see luthe.h)."""

    def __init__(self, ctx, **cfg):
        self.ctx = ctx
        self.cfg = ServerConfig()
        lib().luthe_server_config_default(C.byref(self.cfg))
        for k, v in cfg.items():
            setattr(self.cfg, k, v)
        self._h = C.c_void_p()
        rc = lib().luthe_server_create(ctx.handle, C.byref(self.cfg), C.byref(self._h))
        if rc != 0:
            ctx._chk(rc, "luthe_server_create")
        self._outs = {}
        self.blen = ctx.params.blen

    def _chk(self, rc, what):
        if rc < 0:
            msg = lib().luthe_server_last_error(self._h)
            raise LutheError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def submit(self, op, a, b):
        """Queue IOp `op` on BLWE arrays a, b of shape (nblocks, blen).  Returns the id."""
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b, np.uint64)
        assert a.shape == b.shape and a.ndim == 2 and a.shape[1] == self.blen
        nb = a.shape[0]
        out = np.zeros((1 if op == IOP_CMP_GT else nb, self.blen), np.uint64)
        rid = C.c_uint64()
        self._chk(lib().luthe_server_submit(self._h, op, nb, _vp(a), _vp(b), _vp(out), C.byref(rid)), "submit")
        self._outs[rid.value] = out
        return rid.value

    def poll(self, rid):
        """None while pending, else (out, info dict)."""
        info = RequestInfo()
        rc = self._chk(lib().luthe_server_poll(self._h, rid, C.byref(info)), "poll")
        if rc == 0:
            return None
        return self._outs.pop(rid), {n: getattr(info, n) for n, _ in info._fields_}

    def wait(self, rid, timeout_us=-1):
        info = RequestInfo()
        self._chk(lib().luthe_server_wait(self._h, rid, timeout_us, C.byref(info)), "wait")
        return self._outs.pop(rid), {n: getattr(info, n) for n, _ in info._fields_}

    def flush(self):
        self._chk(lib().luthe_server_flush(self._h), "flush")

    def stats(self, reset=False):
        st = ServerStats()
        self._chk(lib().luthe_server_get_stats(self._h, C.byref(st), 1 if reset else 0), "stats")
        return st.as_dict()

    def close(self):
        if self._h:
            lib().luthe_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- HPU memory images (host only)
def hpu_bsk_order(cut):
    c = np.ascontiguousarray(cut, np.uint32)
    out = np.zeros(1 << int(c.sum()), np.uint32)
    rc = lib().luthe_hpu_bsk_order(_vp(c), c.size, _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_bsk_order")
    return out


def hpu_bsk_to_image(params, layout, bsk_ntt):
    src = np.ascontiguousarray(bsk_ntt, np.uint64)
    out = np.empty(src.size, np.uint64)
    rc = lib().luthe_hpu_bsk_to_image(C.byref(params), C.byref(layout), _vp(src), _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_bsk_to_image")
    return out


def hpu_bsk_from_image(params, layout, image):
    src = np.ascontiguousarray(image, np.uint64)
    out = np.empty(src.size, np.uint64)
    rc = lib().luthe_hpu_bsk_from_image(C.byref(params), C.byref(layout), _vp(src), _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_bsk_from_image")
    return out.reshape(params.bsk_shape())


def hpu_ksk_image_bytes(params, layout):
    return lib().luthe_hpu_ksk_image_bytes(C.byref(params), C.byref(layout))


def hpu_ksk_to_image(params, layout, ksk):
    """One uint8 array per KSK pseudo-channel."""
    k = np.ascontiguousarray(ksk, np.uint32)
    nb = hpu_ksk_image_bytes(params, layout)
    if nb == 0:
        raise LutheError(-1, "invalid KSK layout")
    imgs = [np.zeros(nb, np.uint8) for _ in range(layout.ksk_pc)]
    ptrs = (C.c_void_p * len(imgs))(*[i.ctypes.data for i in imgs])
    rc = lib().luthe_hpu_ksk_to_image(C.byref(params), C.byref(layout), _vp(k), ptrs)
    if rc != 0:
        raise LutheError(rc, "hpu_ksk_to_image")
    return imgs


def hpu_ksk_from_image(params, layout, images):
    imgs = [np.ascontiguousarray(i, np.uint8) for i in images]
    ptrs = (C.c_void_p * len(imgs))(*[i.ctypes.data for i in imgs])
    out = np.zeros(params.ksk_shape(), np.uint32)
    rc = lib().luthe_hpu_ksk_from_image(C.byref(params), C.byref(layout), ptrs, _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_ksk_from_image")
    return out


def hpu_ct_slot_bytes(params, layout):
    return lib().luthe_hpu_ct_slot_bytes(C.byref(params), C.byref(layout))


def hpu_ct_to_image(params, layout, blwe):
    b = np.ascontiguousarray(blwe, np.uint64).reshape(-1, params.blen)
    sb = hpu_ct_slot_bytes(params, layout)
    if sb == 0:
        raise LutheError(-1, "invalid layout")
    out = np.zeros(layout.pem_pc * b.shape[0] * sb, np.uint8)
    rc = lib().luthe_hpu_ct_to_image(C.byref(params), C.byref(layout), _vp(b), b.shape[0], _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_ct_to_image")
    return out


def hpu_ct_from_image(params, layout, image, count):
    img = np.ascontiguousarray(image, np.uint8)
    out = np.zeros((count, params.blen), np.uint64)
    rc = lib().luthe_hpu_ct_from_image(C.byref(params), C.byref(layout), _vp(img), count, _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_ct_from_image")
    return out


def hpu_lut_from_image(params, bodies):
    b = np.ascontiguousarray(bodies, np.uint64).reshape(-1, params.poly_size)
    out = np.zeros((b.shape[0], params.lut_len), np.uint64)
    rc = lib().luthe_hpu_lut_from_image(C.byref(params), _vp(b), b.shape[0], _vp(out))
    if rc != 0:
        raise LutheError(rc, "hpu_lut_from_image")
    return out


def trace_pack(lwe_dim, events):
    n = len(events)
    ev = (TraceEvent * max(1, n))()
    for i, e in enumerate(events):
        for k, v in e.items():
            setattr(ev[i], k, v)
    bits = lib().luthe_trace_record_bits(lwe_dim)
    nw = (bits + 31) // 32
    out = np.zeros(max(1, n) * nw, np.uint32)
    rc = lib().luthe_trace_pack(lwe_dim, C.cast(ev, C.c_void_p), n, _vp(out))
    if rc < 0:
        raise LutheError(rc, "trace_pack")
    return out[:n * nw].reshape(n, nw)


def trace_unpack(lwe_dim, words):
    w = np.ascontiguousarray(words, np.uint32)
    n = w.shape[0]
    ev = (TraceEvent * max(1, n))()
    rc = lib().luthe_trace_unpack(lwe_dim, _vp(w), n, C.cast(ev, C.c_void_p))
    if rc < 0:
        raise LutheError(rc, "trace_unpack")
    return [ev[i].as_dict() for i in range(n)]
