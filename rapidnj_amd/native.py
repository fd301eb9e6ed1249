"""ctypes binding of librapidnj_amd.so (the C ABI declared in include/rapidnj_amd.h).

The library is the product: every distance, count and packing call below runs the HIP kernels.
There is no CPU fallback -- importing this module without the built library raises, and so does
creating a :class:`Context` without a HIP device. Device buffers are torch tensors (allocation
and stream plumbing only).
"""
from __future__ import annotations

import ctypes
import weakref
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librapidnj_amd.so")

RNJ_OK, RNJ_E_INVALID, RNJ_E_HIP, RNJ_E_NOMEM, RNJ_E_NOT_DNA = 0, 1, 2, 3, 4
RNJ_E_ADMISSION, RNJ_E_CANCELLED, RNJ_E_NOT_READY, RNJ_E_NO_DEVICE = 5, 6, 7, 8
JC, K2P = 0, 1
TYPE_AUTO, TYPE_DNA, TYPE_PROTEIN = 0, 1, 2
TRIANGLE, SQUARE = 0, 1
FLAG_RAW = 1
FLAG_EXACT = 2
MIX_K2P, MIX_JC, MIX_NAIVE, MIX_PROTEIN = 0, 1, 2, 3

_vp = ctypes.c_void_p
_u = ctypes.c_uint
_sz = ctypes.c_size_t

# name -> (restype, argtypes); the exported C ABI, in header order
SIGNATURES = {
    "initialiseDNA2_gpu": (None, [_u, _u, _u]),
    "storeDataDNA2_gpu": (None, [_vp, _vp]),
    "computeDistancesDNA2_gpu": (None, [_vp]),
    "initialiseDNA_gpu": (None, [_u, _u, _u]),
    "storeDataDNA_gpu": (None, [_vp, _vp]),
    "computeDistancesDNA_gpu": (None, [_vp, _vp, _vp]),
    "getResultsDNA_gpu": (None, [_vp, _vp, _vp]),
    "initialiseProtein_gpu": (None, [_u, _u, _u]),
    "computeDistancesProtein_gpu": (None, [_vp, _vp, _vp]),
    "getResultsProtein_gpu": (None, [_vp, _vp]),
    "rnj_last_error": (ctypes.c_char_p, []),
    "rnj_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "rnj_api_version": (ctypes.c_int, []),
    "rnj_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "rnj_destroy": (ctypes.c_int, [_vp]),
    "rnj_synchronize": (ctypes.c_int, [_vp]),
    "rnj_tile_config_count": (ctypes.c_int, []),
    "rnj_tile_config_name": (ctypes.c_char_p, [ctypes.c_int]),
    "rnj_set_tile_config": (ctypes.c_int, [_vp, ctypes.c_int]),
    "rnj_get_tile_config_for": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "rnj_set_input_type": (ctypes.c_int, [_vp, ctypes.c_int]),
    "rnj_set_char_rules": (ctypes.c_int, [_vp, ctypes.c_int]),
    "rnj_get_tile_config": (ctypes.c_int, [_vp, _u, _u, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "rnj_alignment_from_chars": (ctypes.c_int, [_vp, _vp, _sz, _u, _u, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(_vp)]),
    "rnj_alignment_from_rows": (ctypes.c_int, [_vp, _vp, _u, _u, ctypes.POINTER(_vp)]),
    "rnj_alignment_from_bitstrings": (ctypes.c_int, [_vp, _vp, _vp, _u, _u, ctypes.POINTER(_vp)]),
    "rnj_alignment_from_fasta": (ctypes.c_int, [_vp, _vp, _sz, ctypes.c_int, ctypes.POINTER(_vp)]),
    "rnj_alignment_from_fasta_file": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "rnj_alignment_from_stockholm": (ctypes.c_int, [_vp, _vp, _sz, ctypes.c_int,
                                                    ctypes.POINTER(_vp)]),
    "rnj_alignment_from_stockholm_file": (ctypes.c_int, [_vp, ctypes.c_char_p,
                                                         ctypes.POINTER(_vp)]),
    "rnj_alignment_name_count": (_u, [_vp]),
    "rnj_alignment_name": (ctypes.c_char_p, [_vp, _u]),
    "rnj_alignment_free": (ctypes.c_int, [_vp]),
    "rnj_alignment_info": (ctypes.c_int, [_vp, ctypes.POINTER(_u), ctypes.POINTER(_u),
                                          ctypes.POINTER(_u), ctypes.POINTER(_sz)]),
    "rnj_alignment_type": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_u)]),
    "rnj_alignment_planes": (ctypes.c_int, [_vp, _vp]),
    "rnj_distances_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _sz,
                                            ctypes.c_int, _vp, _vp]),
    "rnj_distances_device_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                                    _sz, ctypes.c_int, _vp, _vp]),
    "rnj_counts_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rnj_copy_to_host": (ctypes.c_int, [_vp, _vp, _vp, _sz, _vp]),
    "rnj_valu_issue_rates": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]),
    "rnj_valu_mix_rate": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]),
    "rnj_distances_rows": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp]),
    "rnj_distances_rows_lower": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp]),
    "rnj_distances_stream_rows": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "rnj_bootstrap_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _u, ctypes.c_int, _vp,
                                            _sz, _sz, ctypes.c_int, _vp]),
    "rnj_bootstrap_plan_create": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u, _u, ctypes.c_int,
                                                 _vp, _sz, _sz, ctypes.c_int,
                                                 ctypes.POINTER(_vp)]),
    "rnj_bootstrap_plan_run": (ctypes.c_int, [_vp, _vp, _vp]),
    "rnj_plan_tile_config": (ctypes.c_int, [_vp]),
    "rnj_plan_destroy": (ctypes.c_int, [_vp]),
    "rnj_batch_plan_create": (ctypes.c_int, [_vp, _vp, _vp, _u, ctypes.c_int, _vp, _vp,
                                             ctypes.c_int, ctypes.POINTER(_vp)]),
    "rnj_batch_plan_run": (ctypes.c_int, [_vp, _vp]),
    "rnj_batch_plan_tile_config": (ctypes.c_int, [_vp]),
    "rnj_batch_plan_destroy": (ctypes.c_int, [_vp]),
    "rnj_alignment_update_chars": (ctypes.c_int, [_vp, _vp, _vp, _sz, ctypes.c_int, _vp]),
    "rnj_nj_device": (ctypes.c_int, [_vp, _vp, _u, _sz, ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "rnj_nj_rows": (ctypes.c_int, [_vp, _vp, _u, ctypes.c_int, ctypes.POINTER(_vp)]),
    "rnj_nj_alignment": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "rnj_tree_size": (_u, [_vp]),
    "rnj_tree_host_searches": (_u, [_vp]),
    "rnj_tree_joins": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rnj_tree_newick": (ctypes.c_int, [_vp, _vp, _vp, _sz, ctypes.POINTER(_sz)]),
    "rnj_tree_free": (ctypes.c_int, [_vp]),
    "rnj_server_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "rnj_server_submit": (ctypes.c_int, [_vp, _vp, _sz, _u, _u, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "rnj_server_submit_resident": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp,
                                                  ctypes.POINTER(ctypes.c_uint64)]),
    "rnj_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _vp]),
    "rnj_server_poll": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "rnj_server_cancel": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "rnj_server_destroy": (ctypes.c_int, [_vp]),
}


_ROW_SINK = ctypes.CFUNCTYPE(ctypes.c_int, _vp, _u, ctypes.POINTER(ctypes.c_float), _u)


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", _u), ("max_batch_jobs", _u), ("hbm_budget", _sz),
                ("batch_window_us", _u), ("max_batch_work", ctypes.c_ulonglong)]


class JobStats(ctypes.Structure):
    _fields_ = [("queued_s", ctypes.c_double), ("device_s", ctypes.c_double),
                ("total_s", ctypes.c_double), ("batch_id", _u), ("batch_jobs", _u),
                ("upload_s", ctypes.c_double), ("launch_s", ctypes.c_double),
                ("run_s", ctypes.c_double), ("download_s", ctypes.c_double),
                ("launched_at", ctypes.c_double), ("finished_at", ctypes.c_double)]


class RnjError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_status_name(status)}: {message}")
        self.status = status


_lib = None


def lib() -> ctypes.CDLL:
    """Load librapidnj_amd.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime under the same sonames (libamdhip64.so.7,
        # libhsa-runtime64.so.1). Loading torch first makes the dynamic linker bind this library
        # to that one runtime; loaded the other way round a process would hold two HSA runtimes
        # and torch could no longer see the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `make -C rapidnj_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _status_name(status: int) -> str:
    try:
        return lib().rnj_status_string(status).decode()
    except Exception:  # noqa: BLE001 - only used while formatting another error
        return f"status {status}"


def check(status: int) -> None:
    if status != RNJ_OK:
        raise RnjError(status, (lib().rnj_last_error() or b"").decode())


def _model(model) -> int:
    if isinstance(model, str):
        return {"jc": JC, "k2p": K2P, "kim": K2P}[model.lower()]
    return int(model)


def _stream_handle(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


def _host_rows(aln: np.ndarray) -> tuple[np.ndarray, int]:
    """An (n, L) uint8 host alignment and its row pitch in bytes: a view whose rows are
    contiguous (e.g. a column slice of a wider buffer) is passed as it lies, with its pitch."""
    if aln.dtype == np.uint8 and aln.ndim == 2 and aln.strides[1] == 1 and aln.strides[0] >= aln.shape[1]:
        return aln, aln.strides[0]
    a = np.ascontiguousarray(aln, dtype=np.uint8)
    return a, a.shape[1]


@dataclass
class AlignmentInfo:
    n: int
    L: int
    W: int
    bytes: int
    type: int = TYPE_DNA
    nplanes: int = 3


def _type(t) -> int:
    if isinstance(t, str):
        return {"auto": TYPE_AUTO, "dna": TYPE_DNA, "d": TYPE_DNA, "protein": TYPE_PROTEIN,
                "p": TYPE_PROTEIN}[t.lower()]
    return int(t)


class Alignment:
    """An alignment packed into HBM as bit planes (owned by a :class:`Context`): H/L/V for DNA,
    gap bits + residue code bits for protein."""

    def __init__(self, ctx: "Context", handle: int):
        self.ctx = ctx
        self.handle = handle
        ctx._adopt(self)
        n, L, W, b = _u(), _u(), _u(), _sz()
        check(lib().rnj_alignment_info(handle, ctypes.byref(n), ctypes.byref(L), ctypes.byref(W),
                                       ctypes.byref(b)))
        t, P = ctypes.c_int(), _u()
        check(lib().rnj_alignment_type(handle, ctypes.byref(t), ctypes.byref(P)))
        self.info = AlignmentInfo(n.value, L.value, W.value, b.value, t.value, P.value)

    @property
    def is_protein(self) -> bool:
        return self.info.type == TYPE_PROTEIN

    @property
    def n(self) -> int:
        return self.info.n

    @property
    def L(self) -> int:
        return self.info.L

    @property
    def names(self) -> list[bytes]:
        """FASTA record names in file order (empty for alignments not read from FASTA)."""
        k = lib().rnj_alignment_name_count(self.handle)
        return [lib().rnj_alignment_name(self.handle, i) for i in range(k)]

    def update(self, chars, stream=None) -> None:
        """New characters of the same (n, L) shape, packed in place on `stream` (a uint8 numpy
        array on the host, or a uint8 torch tensor on the device). DNA only."""
        if isinstance(chars, np.ndarray):
            a, pitch = _host_rows(chars)
            assert a.shape == (self.n, self.L), "same shape as the alignment"
            check(lib().rnj_alignment_update_chars(self.ctx.handle, self.handle, a.ctypes.data,
                                                   pitch, 0, _stream_handle(stream)))
        else:
            assert chars.dtype.itemsize == 1 and chars.is_cuda and chars.stride(1) == 1
            assert tuple(chars.shape) == (self.n, self.L), "same shape as the alignment"
            check(lib().rnj_alignment_update_chars(self.ctx.handle, self.handle, chars.data_ptr(),
                                                   chars.stride(0), 1, _stream_handle(stream)))

    def planes(self) -> np.ndarray:
        out = np.zeros((self.info.nplanes, self.info.W, self.info.n), dtype=np.uint32)
        check(lib().rnj_alignment_planes(self.handle, out.ctypes.data))
        return out

    def close(self) -> None:
        if self.handle:
            lib().rnj_alignment_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class Context:
    """One HIP device: streams, tile configuration, scratch. Raises without a device."""

    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().rnj_create(device, ctypes.byref(h)))
        self.handle = h.value
        self.device = device
        self._children = weakref.WeakSet()  # alignments, plans and servers made from this context

    def _adopt(self, child):
        self._children.add(child)

    def close(self) -> None:
        """Destroy the context; servers, plans and alignments still open are closed first."""
        if self.handle:
            live = list(self._children)
            for kind in (Server, BootstrapPlan, BatchPlan, Alignment):
                for c in live:
                    if isinstance(c, kind):
                        c.close()
            lib().rnj_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self) -> None:
        check(lib().rnj_synchronize(self.handle))

    def set_tile_config(self, config: int) -> None:
        check(lib().rnj_set_tile_config(self.handle, config))

    def tile_config(self, n: int, L: int, model="k2p") -> int:
        c = ctypes.c_int()
        check(lib().rnj_get_tile_config(self.handle, n, L, _model(model), ctypes.byref(c)))
        return c.value

    def tile_config_for(self, aln: "Alignment", model="k2p") -> int:
        """Configuration the next launch of `aln` with `model` uses."""
        c = ctypes.c_int()
        check(lib().rnj_get_tile_config_for(self.handle, aln.handle, _model(model), ctypes.byref(c)))
        return c.value

    # ---- alignments ---------------------------------------------------------------------------
    def set_input_type(self, type) -> None:
        """"auto" (the reference's discovery on the first sequence), "dna" or "protein"."""
        check(lib().rnj_set_input_type(self.handle, _type(type)))

    def set_char_rules(self, naive: bool) -> None:
        """DNA alignments read from characters afterwards count as the reference's bit engine
        (False, the default) or as its naive `-f` counter (True: RNJ_CHARS_NAIVE)."""
        check(lib().rnj_set_char_rules(self.handle, 1 if naive else 0))

    def alignment(self, aln, skip_type_check: bool = False) -> Alignment:
        """Pack an (n, L) uint8 alignment (numpy on the host, or a uint8 torch tensor on the device)."""
        h = _vp()
        if isinstance(aln, np.ndarray):
            a, pitch = _host_rows(aln)
            n, L = a.shape
            check(lib().rnj_alignment_from_chars(self.handle, a.ctypes.data, pitch, n, L, 0,
                                                 int(skip_type_check), ctypes.byref(h)))
        else:
            import torch
            assert aln.dtype == torch.uint8 and aln.dim() == 2 and aln.is_cuda
            n, L = aln.shape
            pitch = aln.stride(0)
            assert aln.stride(1) == 1
            torch.cuda.current_stream().synchronize()
            check(lib().rnj_alignment_from_chars(self.handle, aln.data_ptr(), pitch, n, L, 1,
                                                 int(skip_type_check), ctypes.byref(h)))
        return Alignment(self, h.value)

    def alignment_from_rows(self, rows: list[bytes]) -> Alignment:
        n, L = len(rows), len(rows[0])
        arr = (ctypes.c_char_p * n)(*rows)
        h = _vp()
        check(lib().rnj_alignment_from_rows(self.handle, ctypes.cast(arr, _vp), n, L,
                                            ctypes.byref(h)))
        return Alignment(self, h.value)

    def alignment_from_bitstrings(self, bits: np.ndarray, gaps: np.ndarray) -> Alignment:
        """From the reference's packed layout: (n, 4*bitStringsCount) uint32 codes and filters."""
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        gaps = np.ascontiguousarray(gaps, dtype=np.uint32)
        n, words = bits.shape
        pb = (_vp * n)(*[bits.ctypes.data + i * bits.strides[0] for i in range(n)])
        pg = (_vp * n)(*[gaps.ctypes.data + i * gaps.strides[0] for i in range(n)])
        h = _vp()
        check(lib().rnj_alignment_from_bitstrings(self.handle, ctypes.cast(pb, _vp),
                                                  ctypes.cast(pg, _vp), n, words // 4,
                                                  ctypes.byref(h)))
        return Alignment(self, h.value)

    def alignment_from_fasta(self, data) -> Alignment:
        """Parse a FASTA image on the device (bytes / uint8 numpy on the host, or a uint8 torch
        tensor on the device) and pack it; record names in Alignment.names."""
        h = _vp()
        if hasattr(data, "data_ptr"):
            check(lib().rnj_alignment_from_fasta(self.handle, data.data_ptr(), data.numel(), 1,
                                                 ctypes.byref(h)))
        else:
            buf = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) \
                else np.ascontiguousarray(data, dtype=np.uint8)
            check(lib().rnj_alignment_from_fasta(self.handle, buf.ctypes.data, buf.size, 0,
                                                 ctypes.byref(h)))
        return Alignment(self, h.value)

    def alignment_from_fasta_file(self, path: str) -> Alignment:
        h = _vp()
        check(lib().rnj_alignment_from_fasta_file(self.handle, os.fsencode(path), ctypes.byref(h)))
        return Alignment(self, h.value)

    def alignment_from_stockholm(self, data) -> Alignment:
        """Parse a single-line Stockholm image on the device (bytes / uint8 numpy on the host, or
        a uint8 torch tensor on the device) and pack it; sequence names in Alignment.names."""
        h = _vp()
        if hasattr(data, "data_ptr"):
            check(lib().rnj_alignment_from_stockholm(self.handle, data.data_ptr(), data.numel(), 1,
                                                     ctypes.byref(h)))
        else:
            buf = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) \
                else np.ascontiguousarray(data, dtype=np.uint8)
            check(lib().rnj_alignment_from_stockholm(self.handle, buf.ctypes.data, buf.size, 0,
                                                     ctypes.byref(h)))
        return Alignment(self, h.value)

    def alignment_from_stockholm_file(self, path: str) -> Alignment:
        h = _vp()
        check(lib().rnj_alignment_from_stockholm_file(self.handle, os.fsencode(path),
                                                      ctypes.byref(h)))
        return Alignment(self, h.value)

    # ---- distances ----------------------------------------------------------------------------
    def distances(self, aln: Alignment, model="k2p", layout: str = "triangle", out=None,
                  raw: bool = False, max_out=None, stream=None, exact: bool = False):
        """Distance matrix into a device tensor (float32): triangle (n(n-1)/2) or square (n, n)."""
        import torch
        n = aln.n
        lay = {"triangle": TRIANGLE, "square": SQUARE}[layout]
        if out is None:
            shape = (n * (n - 1) // 2,) if lay == TRIANGLE else (n, n)
            out = torch.empty(shape, dtype=torch.float32, device=f"cuda:{self.device}")
        ld = out.stride(0) if lay == SQUARE else 0
        check(lib().rnj_distances_device(self.handle, aln.handle, _model(model), lay,
                                         out.data_ptr(), ld,
                                         (FLAG_RAW if raw else 0) | (FLAG_EXACT if exact else 0),
                                         max_out.data_ptr() if max_out is not None else None,
                                         _stream_handle(stream)))
        return out

    def valu_issue_rates(self) -> tuple[float, float]:
        """Measured lane-ops/s of the count loop's bitwise class and of v_bcnt on this device."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().rnj_valu_issue_rates(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def valu_mix_rate(self, mix: str, planes: int = 0) -> float:
        """Measured lane-ops/s of one count loop's instruction sequence ("k2p", "jc", "naive" or
        "protein" with its code planes) on this device: that loop's VALU ceiling."""
        m = {"k2p": MIX_K2P, "jc": MIX_JC, "naive": MIX_NAIVE, "protein": MIX_PROTEIN}[mix]
        r = ctypes.c_double()
        check(lib().rnj_valu_mix_rate(self.handle, m, planes, ctypes.byref(r)))
        return r.value

    def to_host(self, t, stream=None) -> np.ndarray:
        """A contiguous device tensor as a numpy array, copied through the library's pinned
        staging chunks (rnj_copy_to_host) after the work enqueued on `stream`."""
        import torch
        assert t.is_cuda and t.is_contiguous()
        out = np.empty(tuple(t.shape), dtype=torch.empty(0, dtype=t.dtype).numpy().dtype)
        check(lib().rnj_copy_to_host(self.handle, out.ctypes.data, t.data_ptr(),
                                     t.numel() * t.element_size(), _stream_handle(stream)))
        return out

    def counts(self, aln: Alignment, stream=None):
        """(ts, tv, valid) upper-triangle uint32 device tensors (as int32 storage)."""
        import torch
        m = aln.n * (aln.n - 1) // 2
        buf = torch.empty((3, max(m, 1)), dtype=torch.int32, device=f"cuda:{self.device}")
        check(lib().rnj_counts_device(self.handle, aln.handle, buf[0].data_ptr(),
                                      buf[1].data_ptr(), buf[2].data_ptr(),
                                      _stream_handle(stream)))
        return buf[0, :m], buf[1, :m], buf[2, :m]

    def distance_rows(self, aln: Alignment, model="k2p") -> tuple[np.ndarray, float]:
        """Post-processed square matrix handed to host rows (the in-memory NJ hand-off)."""
        n = aln.n
        mat = np.empty((n, n), dtype=np.float32)
        rows = (_vp * n)(*[mat.ctypes.data + i * mat.strides[0] for i in range(n)])
        mx = ctypes.c_float()
        check(lib().rnj_distances_rows(self.handle, aln.handle, _model(model),
                                       ctypes.cast(rows, _vp), ctypes.byref(mx)))
        return mat, mx.value

    def distance_rows_lower(self, aln: Alignment, model="k2p") -> tuple[list[np.ndarray], float]:
        """Lower-triangle hand-off: row i holds the i + 1 values up to the diagonal."""
        n = aln.n
        out = [np.empty(i + 1, dtype=np.float32) for i in range(n)]
        rows = (_vp * n)(*[r.ctypes.data for r in out])
        mx = ctypes.c_float()
        check(lib().rnj_distances_rows_lower(self.handle, aln.handle, _model(model),
                                             ctypes.cast(rows, _vp), ctypes.byref(mx)))
        return out, mx.value

    def distance_stream_rows(self, aln: Alignment, model="k2p") -> tuple[np.ndarray, float]:
        """The disk-matrix row stream (rnj_distances_stream_rows): every post-processed row as
        the library hands it to the sink, in order, gathered into an (n, n) array."""
        n = aln.n
        mat = np.empty((n, n), dtype=np.float32)
        seen = []

        def sink(user, row, values, length):
            mat[row] = np.ctypeslib.as_array(values, shape=(length,))
            seen.append(row)
            return 0

        cb = _ROW_SINK(sink)
        mx = ctypes.c_float()
        check(lib().rnj_distances_stream_rows(self.handle, aln.handle, _model(model),
                                              ctypes.cast(cb, _vp), None, ctypes.byref(mx)))
        assert seen == list(range(n)), "rows arrive once each, in order"
        return mat, mx.value

    def bootstrap(self, aln: Alignment, weights: np.ndarray, model="k2p",
                  layout: str = "triangle", out=None, raw: bool = False, stream=None,
                  exact: bool = False):
        """R replicate matrices from (R, L) uint16 column weights, in one launch."""
        import torch
        w = np.ascontiguousarray(weights, dtype=np.uint16)
        R, L = w.shape
        assert L == aln.L, "one weight per alignment column"
        n = aln.n
        lay = {"triangle": TRIANGLE, "square": SQUARE}[layout]
        if out is None:
            shape = (R, n * (n - 1) // 2) if lay == TRIANGLE else (R, n, n)
            out = torch.empty(shape, dtype=torch.float32, device=f"cuda:{self.device}")
        mstride = out.stride(0)
        ld = out.stride(1) if lay == SQUARE else 0
        check(lib().rnj_bootstrap_device(self.handle, aln.handle, _model(model), w.ctypes.data, R,
                                         lay, out.data_ptr(), mstride, ld,
                                         (FLAG_RAW if raw else 0) | (FLAG_EXACT if exact else 0),
                                         _stream_handle(stream)))
        return out

    def bootstrap_plan(self, aln: Alignment, R: int, model="k2p", layout: str = "triangle",
                       out=None, max_length: int = 0, raw: bool = False,
                       exact: bool = False) -> "BootstrapPlan":
        """A captured (HIP graph) batched-bootstrap launch writing R matrices into `out`."""
        return BootstrapPlan(self, aln, R, model, layout, out, max_length, raw, exact)

    def batch_plan(self, alns: list, models, layout: str = "triangle", outs=None,
                   raw: bool = False, exact: bool = False) -> "BatchPlan":
        """Several alignments (any n, L) in one captured tile launch; outs[k] receives matrix k."""
        return BatchPlan(self, alns, models, layout, outs, raw, exact)

    def nj(self, matrix, negative_branches: bool = False, stream=None) -> Tree:
        """rapidNJ on a square float32 device matrix (torch tensor (n, ld>=n) on this device)."""
        import torch
        assert matrix.dtype == torch.float32 and matrix.is_cuda and matrix.dim() == 2
        assert matrix.stride(1) == 1
        n = matrix.shape[0]
        h = _vp()
        check(lib().rnj_nj_device(self.handle, matrix.data_ptr(), n, matrix.stride(0),
                                  int(negative_branches), _stream_handle(stream), ctypes.byref(h)))
        return Tree(h.value)

    def nj_rows(self, matrix: np.ndarray, negative_branches: bool = False) -> Tree:
        """rapidNJ on a host matrix handed over as row pointers (rnj_nj_rows)."""
        m = np.ascontiguousarray(matrix, dtype=np.float32)
        n = m.shape[0]
        rows = (_vp * n)(*[m.ctypes.data + i * m.strides[0] for i in range(n)])
        h = _vp()
        check(lib().rnj_nj_rows(self.handle, ctypes.cast(rows, _vp), n, int(negative_branches),
                                ctypes.byref(h)))
        return Tree(h.value)

    def nj_alignment(self, aln: Alignment, model="k2p", negative_branches: bool = False) -> Tree:
        """Distances and neighbour joining with the matrix kept in HBM."""
        h = _vp()
        check(lib().rnj_nj_alignment(self.handle, aln.handle, _model(model),
                                     int(negative_branches), ctypes.byref(h)))
        return Tree(h.value)

    def server(self, slots: int = 4, max_batch_jobs: int = 64, hbm_budget: int = 0,
               batch_window_us: int = 0, max_batch_work: int = 0) -> "Server":
        return Server(self, slots, max_batch_jobs, hbm_budget, batch_window_us, max_batch_work)


class Tree:
    """A neighbour-joining tree built on the device (rnj_nj_device / rnj_nj_alignment)."""

    def __init__(self, handle: int):
        self.handle = handle
        self.n = lib().rnj_tree_size(handle)
        self.host_searches = lib().rnj_tree_host_searches(handle)

    def joins(self):
        """(left, right, left_len, right_len, last (2,), last_dist) as polytree receives them."""
        k = max(self.n - 2, 1)
        left, right = np.zeros(k, dtype=np.int32), np.zeros(k, dtype=np.int32)
        ll, rl = np.zeros(k, dtype=np.float64), np.zeros(k, dtype=np.float64)
        last = np.zeros(2, dtype=np.int32)
        ld = ctypes.c_float()
        check(lib().rnj_tree_joins(self.handle, left.ctypes.data, right.ctypes.data,
                                   ll.ctypes.data, rl.ctypes.data, last.ctypes.data,
                                   ctypes.byref(ld)))
        m = self.n - 2
        return left[:m], right[:m], ll[:m], rl[:m], last, ld.value

    def newick(self, names=None) -> bytes:
        arr = None
        if names is not None:
            arr = (ctypes.c_char_p * self.n)(*[bytes(x) for x in names])
        length = _sz()
        check(lib().rnj_tree_newick(self.handle, arr, None, 0, ctypes.byref(length)))
        buf = ctypes.create_string_buffer(length.value + 1)
        check(lib().rnj_tree_newick(self.handle, arr, buf, length.value + 1, ctypes.byref(length)))
        return buf.raw[: length.value]

    def close(self) -> None:
        if self.handle:
            lib().rnj_tree_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class BootstrapPlan:
    """rnj_bootstrap_plan_*: one graph per (alignment, model, R, layout, output)."""

    def __init__(self, ctx: Context, aln: Alignment, R: int, model, layout: str, out,
                 max_length: int, raw: bool, exact: bool):
        import torch
        self.ctx, self.aln, self.R = ctx, aln, R
        n = aln.n
        lay = {"triangle": TRIANGLE, "square": SQUARE}[layout]
        if out is None:
            shape = (R, n * (n - 1) // 2) if lay == TRIANGLE else (R, n, n)
            out = torch.empty(shape, dtype=torch.float32, device=f"cuda:{ctx.device}")
        self.out = out
        ld = out.stride(1) if lay == SQUARE else 0
        h = _vp()
        check(lib().rnj_bootstrap_plan_create(ctx.handle, aln.handle, _model(model), R,
                                              max_length, lay, out.data_ptr(), out.stride(0), ld,
                                              (FLAG_RAW if raw else 0) |
                                              (FLAG_EXACT if exact else 0), ctypes.byref(h)))
        self.handle = h.value
        ctx._adopt(self)

    @property
    def tile_config(self) -> int:
        return lib().rnj_plan_tile_config(self.handle)

    def run(self, weights: np.ndarray, stream=None):
        w = np.ascontiguousarray(weights, dtype=np.uint16)
        assert w.shape == (self.R, self.aln.L), "one weight per replicate and column"
        check(lib().rnj_bootstrap_plan_run(self.handle, w.ctypes.data, _stream_handle(stream)))
        return self.out

    def close(self) -> None:
        if self.handle:
            lib().rnj_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class BatchPlan:
    """rnj_batch_plan_*: the matrices of several alignments as one captured launch."""

    def __init__(self, ctx: Context, alns: list, models, layout: str, outs, raw: bool,
                 exact: bool):
        import torch
        self.ctx, self.alns = ctx, list(alns)
        k = len(self.alns)
        models = [models] * k if isinstance(models, (str, int)) else list(models)
        lay = {"triangle": TRIANGLE, "square": SQUARE}[layout]
        if outs is None:
            outs = [torch.empty((a.n * (a.n - 1) // 2,) if lay == TRIANGLE else (a.n, a.n),
                                dtype=torch.float32, device=f"cuda:{ctx.device}")
                    for a in self.alns]
        self.outs = list(outs)
        ha = (_vp * k)(*[a.handle for a in self.alns])
        hm = (ctypes.c_int * k)(*[_model(m) for m in models])
        ho = (_vp * k)(*[o.data_ptr() for o in self.outs])
        hl = (_sz * k)(*[o.stride(0) if lay == SQUARE else 0 for o in self.outs])
        h = _vp()
        check(lib().rnj_batch_plan_create(ctx.handle, ctypes.cast(ha, _vp), ctypes.cast(hm, _vp), k,
                                          lay, ctypes.cast(ho, _vp), ctypes.cast(hl, _vp),
                                          (FLAG_RAW if raw else 0) | (FLAG_EXACT if exact else 0),
                                          ctypes.byref(h)))
        self.handle = h.value
        ctx._adopt(self)

    @property
    def tile_config(self) -> int:
        return lib().rnj_batch_plan_tile_config(self.handle)

    def run(self, stream=None) -> list:
        check(lib().rnj_batch_plan_run(self.handle, _stream_handle(stream)))
        return self.outs

    def close(self) -> None:
        if self.handle:
            lib().rnj_batch_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Served jobs: admission by HBM budget, batching, overlap of batches on the device."""

    def __init__(self, ctx: Context, slots: int, max_batch_jobs: int, hbm_budget: int,
                 batch_window_us: int = 0, max_batch_work: int = 0):
        self.ctx = ctx
        cfg = ServerConfig(slots, max_batch_jobs, hbm_budget, batch_window_us, max_batch_work)
        h = _vp()
        check(lib().rnj_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        ctx._adopt(self)

    def submit(self, aln: np.ndarray, model="k2p") -> int:
        a, pitch = _host_rows(aln)
        n, L = a.shape
        t = ctypes.c_uint64()
        check(lib().rnj_server_submit(self.handle, a.ctypes.data, pitch, n, L, _model(model),
                                      ctypes.byref(t)))
        return t.value

    def submit_resident(self, aln: "Alignment", model="k2p", out=None) -> int:
        """A job on an alignment already in HBM; `out` (a float32 device tensor of n(n-1)/2) receives
        the triangle, else the server holds it until wait() copies it to the host."""
        if out is not None:
            assert out.is_cuda and out.is_contiguous() and out.numel() >= aln.n * (aln.n - 1) // 2
        t = ctypes.c_uint64()
        check(lib().rnj_server_submit_resident(self.handle, aln.handle, _model(model),
                                               out.data_ptr() if out is not None else None,
                                               ctypes.byref(t)))
        return t.value

    def poll(self, ticket: int) -> bool:
        rc = lib().rnj_server_poll(self.handle, ticket)
        if rc == RNJ_E_NOT_READY:
            return False
        check(rc)
        return True

    def wait(self, ticket: int, n: int, copy: bool = True):
        """(triangle on the host, or None without `copy`; the job's JobStats)."""
        out = np.empty(max(n * (n - 1) // 2, 1), dtype=np.float32) if copy else None
        st = JobStats()
        check(lib().rnj_server_wait(self.handle, ticket, out.ctypes.data if copy else None,
                                    ctypes.byref(st)))
        return (out[: n * (n - 1) // 2] if copy else None), st

    def cancel(self, ticket: int) -> None:
        check(lib().rnj_server_cancel(self.handle, ticket))

    def close(self) -> None:
        if self.handle:
            lib().rnj_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def tile_configs() -> list[str]:
    return [lib().rnj_tile_config_name(i).decode() for i in range(lib().rnj_tile_config_count())]
