"""Python binding of liboz2.so, the MI355X Ozaki Scheme II DGEMM/ZGEMM library.

Mirrors the reference's C++ class oz2::Ozaki2 (src/ozaki2_gpu.hpp:678-786)
over the C ABI in include/oz2/oz2_c.h: same member names, argument meaning
and error classes (std::invalid_argument -> OZ2InvalidArgument, a ValueError;
std::runtime_error -> OZ2RuntimeError; ...).  Matrices are column-major
device buffers passed as torch tensors or raw device addresses (int).

There is no CPU fallback: if liboz2.so is missing or cannot be loaded,
importing this package raises.
"""
import ctypes
import os

try:  # load torch's HIP runtime first so liboz2.so binds to the same one
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - the library itself does not need torch
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liboz2.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"oz2: {LIB_PATH} is not built (run __graft_entry__.build() or make -C os2-xe_amd)")
_lib = ctypes.CDLL(LIB_PATH)

N, T, C = 0, 1, 2
NOALLOC_SCRATCH_BYTES = 64 << 20  # OZ2_NOALLOC_SCRATCH_BYTES in oz2_c.h
_TRANS = {"N": 0, "T": 1, "C": 2, 0: 0, 1: 1, 2: 2}

_i64, _int, _dbl, _vp, _sz = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t


class StageTimes(ctypes.Structure):
    _fields_ = [("scale_ms", _dbl), ("split_ms", _dbl), ("gemm_ms", _dbl), ("crt_ms", _dbl)]

    def total_ms(self):
        return self.scale_ms + self.split_ms + self.gemm_ms + self.crt_ms

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TileInfo(ctypes.Structure):
    _fields_ = [("mb", _i64), ("nb", _i64), ("pm", _i64), ("pn", _i64), ("k_chunks", _i64),
                ("workspace_bytes", _sz)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("batched_gemm", _int), ("pad_lda", _int), ("pow2_tiles", _int), ("gemm_tile", _int),
                ("overlap_blocks", _int)]


class OZ2Error(Exception):
    pass


class OZ2InvalidArgument(OZ2Error, ValueError):
    pass


class OZ2RuntimeError(OZ2Error, RuntimeError):
    pass


class OZ2LogicError(OZ2Error, RuntimeError):
    pass


class OZ2OverflowError(OZ2Error, OverflowError):
    pass


_ERRORS = {1: OZ2InvalidArgument, 2: OZ2RuntimeError, 3: OZ2LogicError, 4: OZ2OverflowError}

_P = ctypes.POINTER
_lib.oz2_create.argtypes = [_vp, _P(_vp)]
_lib.oz2_destroy.argtypes = [_vp]
_lib.oz2_set_stream.argtypes = [_vp, _vp]
_lib.oz2_dgemm.argtypes = [_vp, _int, _int, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64, _int,
                           _P(_vp), _int, _vp]
_lib.oz2_zgemm.argtypes = [_vp, _int, _int, _i64, _i64, _i64, _P(_dbl), _vp, _i64, _vp, _i64, _P(_dbl), _vp, _i64,
                           _int, _P(_vp), _int, _vp]
_lib.oz2_dgemm_noalloc.argtypes = _lib.oz2_dgemm.argtypes
_lib.oz2_precision_bits.argtypes = [_i64, _int, _P(_int)]
_lib.oz2_precision_bits_complex.argtypes = [_i64, _int, _P(_int)]
_lib.oz2_moduli_for_bits.argtypes = [_i64, _int]
_lib.oz2_set_workspace_limit.argtypes = [_vp, _sz]
_lib.oz2_set_resident_hint.argtypes = [_vp, _sz]
_lib.oz2_release_workspace.argtypes = [_vp]
_lib.oz2_reserve_workspace.argtypes = [_vp, _sz]
_lib.oz2_workspace_query.argtypes = [_vp, _int, _int, _i64, _i64, _i64, _int, _int, _P(_sz)]
_lib.oz2_stage_times.argtypes = [_vp, _P(StageTimes)]
_lib.oz2_last_tiling.argtypes = [_vp, _P(TileInfo)]
_lib.oz2_set_tuning.argtypes = [_vp, _P(Tuning)]
_lib.oz2_get_tuning.argtypes = [_vp, _P(Tuning)]
_lib.oz2_set_profiling.argtypes = [_vp, _int]
_lib.oz2_moduli.argtypes = [_int, _P(ctypes.c_uint16)]
_lib.oz2_profile_begin.argtypes = [_vp]
_lib.oz2_profile_end.argtypes = [_vp, _P(StageTimes), _P(_i64)]
_lib.oz2_unit_split.argtypes = [_vp, _vp, _i64, _i64, _i64, _int, _int, _int, _vp, _int, _int, _vp, _i64, _i64]
_lib.oz2_unit_crt.argtypes = [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _int, _dbl, _dbl, _vp, _i64]
_lib.oz2_unit_gemm.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _int, _int, _vp, _i64]
_lib.oz2_ozaki1_dgemm.argtypes = [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int]


class GemmDesc(ctypes.Structure):
    """oz2_gemm_desc_t: one DGEMM of a batch (device addresses as ints)"""
    _fields_ = [("ta", _int), ("tb", _int), ("m", _i64), ("n", _i64), ("k", _i64), ("alpha", _dbl), ("A", _vp),
                ("lda", _i64), ("B", _vp), ("ldb", _i64), ("beta", _dbl), ("C", _vp), ("ldc", _i64),
                ("num_moduli", _int)]


_lib.oz2_batch_create.argtypes = [_vp, _int, _P(GemmDesc), _int, _P(_int), _int, _P(_vp)]
_lib.oz2_batch_run.argtypes = [_vp]
_lib.oz2_batch_reselect.argtypes = [_vp, _P(_int)]
_lib.oz2_batch_moduli.argtypes = [_vp, _P(_int)]
_lib.oz2_batch_tiles.argtypes = [_vp, _P(_int)]
_lib.oz2_batch_info.argtypes = [_vp, _P(_sz), _P(_int), _P(_int)]
_lib.oz2_batch_destroy.argtypes = [_vp]
_lib.oz2_batch_autotune.argtypes = [_vp, _int, _P(GemmDesc), _int, _int, _P(_int), _P(_int), _P(_dbl)]
BATCH_GRAPH, BATCH_SELECT, BATCH_PERSISTENT, BATCH_PROBLEM_ORDER = 1, 2, 4, 8
SEL_FEATURES, SEL_OUTPUTS, SEL_PARAMS = 48, 23, 8791
_flt = ctypes.c_float
_lib.oz2_selector_features.argtypes = [_vp, _int, _P(GemmDesc), _int, _P(_flt)]
_lib.oz2_selector_choose.argtypes = [_vp, _int, _P(GemmDesc), _int, _P(_flt), _P(_int), _P(_int)]
_lib.oz2_selector_set_weights.argtypes = [_vp, _P(_flt), _sz, _flt]
_lib.oz2_selector_get_weights.argtypes = [_vp, _P(_flt), _sz, _P(_flt), _P(_sz)]
_lib.oz2_mlp_forward.argtypes = [_vp, _vp, _int, _vp, _vp, _vp, _vp, _vp]
_lib.oz2_mlp_backward.argtypes = [_vp, _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _flt, _vp, _vp, _vp, _vp, _vp]
_lib.oz2_mlp_adam.argtypes = [_vp, _vp, _vp, _vp, _vp, _int, _flt, _flt, _flt, _flt, _flt]
_lib.oz2_last_error.restype = ctypes.c_char_p
_lib.oz2_version.restype = ctypes.c_char_p

EXPORTS = [
    "oz2_create", "oz2_destroy", "oz2_set_stream", "oz2_dgemm", "oz2_zgemm", "oz2_dgemm_noalloc", "oz2_precision_bits",
    "oz2_precision_bits_complex", "oz2_moduli_for_bits", "oz2_set_workspace_limit", "oz2_set_resident_hint",
    "oz2_release_workspace", "oz2_reserve_workspace", "oz2_workspace_query", "oz2_stage_times", "oz2_last_tiling",
    "oz2_set_tuning", "oz2_get_tuning", "oz2_set_profiling", "oz2_profile_begin", "oz2_profile_end", "oz2_moduli",
    "oz2_unit_split", "oz2_unit_crt", "oz2_unit_gemm", "oz2_ozaki1_dgemm", "oz2_batch_create", "oz2_batch_run", "oz2_batch_reselect", "oz2_batch_moduli", "oz2_batch_tiles",
    "oz2_batch_info", "oz2_batch_destroy", "oz2_batch_autotune", "oz2_selector_features", "oz2_selector_choose",
    "oz2_selector_set_weights", "oz2_selector_get_weights", "oz2_mlp_forward", "oz2_mlp_backward", "oz2_mlp_adam",
    "oz2_last_error", "oz2_version",
]


def lib():
    return _lib


def _check(rc):
    if rc != 0:
        msg = _lib.oz2_last_error().decode()
        raise _ERRORS.get(rc, OZ2Error)(msg)


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if torch is not None and isinstance(x, torch.Tensor):
        if not x.is_cuda:
            raise OZ2InvalidArgument("oz2: tensors must live on the GPU")
        return x.data_ptr()
    raise TypeError(f"oz2: expected a device tensor or address, got {type(x)}")


def moduli(num_moduli):
    out = (ctypes.c_uint16 * 24)()
    _check(_lib.oz2_moduli(num_moduli, out))
    return [int(out[i]) for i in range(num_moduli)]


def unit_split(X, ldx, rows, k, kcontig, ex, num_moduli, tk, planes, rows_pad, kp, cplx=False, conj=False,
               stream=None):
    """Kernel-level split (the reference's detail::launch_split in its unit test)."""
    _check(_lib.oz2_unit_split(stream, _ptr(X), ldx, rows, k, int(kcontig), int(cplx), int(conj), _ptr(ex), num_moduli,
                               tk, _ptr(planes), rows_pad, kp))


def unit_crt(R, ldr, mb, nb, D, E, num_moduli, alpha, beta, C, ldc, stream=None):
    """Kernel-level CRT + epilogue (the reference's detail::launch_crt in its unit test)."""
    _check(_lib.oz2_unit_crt(stream, _ptr(R), ldr, mb, nb, _ptr(D), _ptr(E), num_moduli, float(alpha), float(beta),
                             _ptr(C), ldc))


def unit_gemm(A, B, mpad, npad, kp, num_moduli, tile, R, ldr, stream=None):
    """Kernel-level INT8 GEMM with the residue epilogue (the reference's gemm_batch call)."""
    _check(_lib.oz2_unit_gemm(stream, _ptr(A), _ptr(B), mpad, npad, kp, num_moduli, int(tile), _ptr(R), ldr))


def _descs(problems):
    n = len(problems)
    arr = (GemmDesc * max(1, n))()
    for i, p in enumerate(problems):
        arr[i] = GemmDesc(_TRANS.get(p.get("ta", "N"), -1), _TRANS.get(p.get("tb", "N"), -1), p["m"], p["n"],
                          p["k"], float(p.get("alpha", 1.0)), _ptr(p["A"]), p["lda"], _ptr(p["B"]), p["ldb"],
                          float(p.get("beta", 0.0)), _ptr(p.get("C")), p.get("ldc", max(1, p["m"])),
                          int(p.get("num_moduli", 0)))
    return arr


class Batch:
    """A plan for many DGEMMs (oz2_batch_create).  problems: dicts with keys ta, tb, m, n, k, alpha, A, lda,
    B, ldb, beta, C, ldc and optional num_moduli (<= 0 or absent: chosen by the plan for `bits`).
    The plan keeps the tensors it was given alive; run() recomputes from their current contents with the
    moduli chosen at creation; reselect() chooses again from the current contents."""

    def __init__(self, oz, problems, bits=53, tiles=None, graph=True, select=False, persistent=False,
                 largest_first=True):
        self._oz = oz
        self._keep = [(p["A"], p["B"], p["C"]) for p in problems]
        n = len(problems)
        arr = _descs(problems)
        t = None
        if tiles is not None:
            t = (_int * max(1, n))(*tiles)
        flags = ((BATCH_GRAPH if graph else 0) | (BATCH_SELECT if select else 0) |
                 (BATCH_PERSISTENT if persistent else 0) | (0 if largest_first else BATCH_PROBLEM_ORDER))
        self._b = _vp()
        _check(_lib.oz2_batch_create(oz.handle, n, arr, int(bits), t, flags, ctypes.byref(self._b)))
        self.count = n

    def run(self):
        _check(_lib.oz2_batch_run(self._b))

    def reselect(self):
        """re-run the selector on the buffers' current contents (oz2_batch_reselect); returns how many
        problems' num_moduli changed"""
        changed = _int(0)
        _check(_lib.oz2_batch_reselect(self._b, ctypes.byref(changed)))
        return changed.value

    def moduli(self):
        out = (_int * max(1, self.count))()
        _check(_lib.oz2_batch_moduli(self._b, out))
        return list(out)[:self.count]

    def tiles(self):
        out = (_int * max(1, self.count))()
        _check(_lib.oz2_batch_tiles(self._b, out))
        return list(out)[:self.count]

    def info(self):
        ws, launches, graph = _sz(0), _int(0), _int(0)
        _check(_lib.oz2_batch_info(self._b, ctypes.byref(ws), ctypes.byref(launches), ctypes.byref(graph)))
        return {"workspace_bytes": ws.value, "launches": launches.value, "graph": bool(graph.value)}

    def close(self):
        if getattr(self, "_b", None):
            _lib.oz2_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version():
    return _lib.oz2_version().decode()


class Ozaki2:
    """oz2::Ozaki2 bound to a HIP stream (torch stream, raw hipStream_t address, or None = default)."""

    def __init__(self, stream=None):
        self._h = _vp()
        _check(_lib.oz2_create(self._stream_ptr(stream), ctypes.byref(self._h)))

    @staticmethod
    def _stream_ptr(stream):
        if stream is None:
            return None
        if isinstance(stream, int):
            return stream
        return stream.cuda_stream  # torch.cuda.Stream

    def close(self):
        if self._h:
            _lib.oz2_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h.value

    def set_stream(self, stream):
        _check(_lib.oz2_set_stream(self._h, self._stream_ptr(stream)))

    # -- the two GEMMs ----------------------------------------------------------------
    def gemm(self, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, num_moduli, deps=(), done=None):
        dep = (_vp * max(1, len(deps)))(*[int(d) for d in deps])
        _check(_lib.oz2_dgemm(self._h, _TRANS.get(ta, -1), _TRANS.get(tb, -1), m, n, k, float(alpha), _ptr(A), lda, _ptr(B), ldb,
                              float(beta), _ptr(C), ldc, num_moduli, dep, len(deps), done))

    def gemm_noalloc(self, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, num_moduli, deps=(), done=None):
        """gemm() with every intermediate in the library's fixed per-GPU scratch (oz2_dgemm_noalloc)"""
        dep = (_vp * max(1, len(deps)))(*[int(d) for d in deps])
        _check(_lib.oz2_dgemm_noalloc(self._h, _TRANS.get(ta, -1), _TRANS.get(tb, -1), m, n, k, float(alpha), _ptr(A), lda,
                                      _ptr(B), ldb, float(beta), _ptr(C), ldc, num_moduli, dep, len(deps), done))

    def ozaki1(self, m, n, k, A, lda, B, ldb, C, ldc, L):
        """Ozaki Scheme I with L slices (oz2_ozaki1_dgemm): C = A B, op N, alpha 1, beta 0"""
        _check(_lib.oz2_ozaki1_dgemm(self._h, m, n, k, _ptr(A), lda, _ptr(B), ldb, _ptr(C), ldc, int(L)))

    def zgemm(self, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, num_moduli, deps=(), done=None):
        al = (_dbl * 2)(complex(alpha).real, complex(alpha).imag)
        be = (_dbl * 2)(complex(beta).real, complex(beta).imag)
        dep = (_vp * max(1, len(deps)))(*[int(d) for d in deps])
        _check(_lib.oz2_zgemm(self._h, _TRANS.get(ta, -1), _TRANS.get(tb, -1), m, n, k, al, _ptr(A), lda, _ptr(B), ldb, be, _ptr(C),
                              ldc, num_moduli, dep, len(deps), done))

    # -- plan queries -------------------------------------------------------------------
    @staticmethod
    def precision_bits(k, num_moduli):
        err = _int(0)
        tk = _lib.oz2_precision_bits(k, num_moduli, ctypes.byref(err))
        _check(err.value)
        return tk

    @staticmethod
    def precision_bits_complex(k, num_moduli):
        err = _int(0)
        tk = _lib.oz2_precision_bits_complex(k, num_moduli, ctypes.byref(err))
        _check(err.value)
        return tk

    @staticmethod
    def moduli_for_bits(k, bits):
        return _lib.oz2_moduli_for_bits(k, bits)

    # -- memory ---------------------------------------------------------------------------
    def set_workspace_limit(self, nbytes):
        _check(_lib.oz2_set_workspace_limit(self._h, nbytes))

    def set_resident_hint(self, nbytes):
        _check(_lib.oz2_set_resident_hint(self._h, nbytes))

    def release_workspace(self):
        _check(_lib.oz2_release_workspace(self._h))

    def reserve_workspace(self, nbytes):
        _check(_lib.oz2_reserve_workspace(self._h, nbytes))

    def workspace_query(self, ta, tb, m, n, k, num_moduli, cplx=False):
        out = _sz(0)
        _check(_lib.oz2_workspace_query(self._h, _TRANS.get(ta, -1), _TRANS.get(tb, -1), m, n, k, num_moduli, int(cplx),
                                        ctypes.byref(out)))
        return out.value

    # -- introspection -------------------------------------------------------------------
    def stage_times(self):
        st = StageTimes()
        _check(_lib.oz2_stage_times(self._h, ctypes.byref(st)))
        return st

    def last_tiling(self):
        ti = TileInfo()
        _check(_lib.oz2_last_tiling(self._h, ctypes.byref(ti)))
        return ti

    def set_tuning(self, batched_gemm=True, pad_lda=True, pow2_tiles=True, gemm_tile=-1, overlap_blocks=-1):
        t = Tuning(int(batched_gemm), int(pad_lda), int(pow2_tiles), int(gemm_tile), int(overlap_blocks))
        _check(_lib.oz2_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self):
        t = Tuning()
        _check(_lib.oz2_get_tuning(self._h, ctypes.byref(t)))
        return {"batched_gemm": bool(t.batched_gemm), "pad_lda": bool(t.pad_lda), "pow2_tiles": bool(t.pow2_tiles),
                "gemm_tile": t.gemm_tile, "overlap_blocks": t.overlap_blocks}

    def profile_begin(self):
        _check(_lib.oz2_profile_begin(self._h))

    def profile_end(self):
        """(summed StageTimes, number of calls) since profile_begin()"""
        st = StageTimes()
        calls = _i64(0)
        _check(_lib.oz2_profile_end(self._h, ctypes.byref(st), ctypes.byref(calls)))
        return st, calls.value

    def set_profiling(self, on):
        _check(_lib.oz2_set_profiling(self._h, int(bool(on))))

    # -- learned moduli selector ------------------------------------------------------
    def selector_features(self, problems, bits=53):
        """list of problem dicts (as for Batch, C not needed) -> numpy array (n, 48)"""
        import numpy as np
        n = len(problems)
        out = (_flt * max(1, n * SEL_FEATURES))()
        _check(_lib.oz2_selector_features(self._h, n, _descs(problems), int(bits), out))
        return np.ctypeslib.as_array(out)[:n * SEL_FEATURES].reshape(n, SEL_FEATURES).copy()

    def selector_choose(self, problems, bits=53):
        """-> (s per problem, probabilities (n, 23), learned flags)"""
        import numpy as np
        n = len(problems)
        pr = (_flt * max(1, n * SEL_OUTPUTS))()
        s = (_int * max(1, n))()
        lf = (_int * max(1, n))()
        _check(_lib.oz2_selector_choose(self._h, n, _descs(problems), int(bits), pr, s, lf))
        return (list(s)[:n], np.ctypeslib.as_array(pr)[:n * SEL_OUTPUTS].reshape(n, SEL_OUTPUTS).copy(),
                [bool(x) for x in list(lf)[:n]])

    def selector_set_weights(self, weights, threshold=0.9):
        import numpy as np
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        _check(_lib.oz2_selector_set_weights(self._h, w.ctypes.data_as(_P(_flt)), w.size, float(threshold)))

    def selector_weights(self):
        """-> (weights as float32 numpy array or None, threshold)"""
        import numpy as np
        w = np.zeros(SEL_PARAMS, dtype=np.float32)
        tau, cnt = _flt(0), _sz(0)
        _check(_lib.oz2_selector_get_weights(self._h, w.ctypes.data_as(_P(_flt)), w.size, ctypes.byref(tau),
                                             ctypes.byref(cnt)))
        return (w if cnt.value else None), tau.value

    def mlp_forward(self, X, nb, W, H1=None, H2=None, Z=None, P=None):
        _check(_lib.oz2_mlp_forward(self._h, _ptr(X), nb, _ptr(W), _ptr(H1), _ptr(H2), _ptr(Z), _ptr(P)))

    def mlp_backward(self, X, nb, W, H1, H2, Z, Y, M, inv_norm, D1, D2, D3, G, loss=None):
        _check(_lib.oz2_mlp_backward(self._h, _ptr(X), nb, _ptr(W), _ptr(H1), _ptr(H2), _ptr(Z), _ptr(Y), _ptr(M),
                                     float(inv_norm), _ptr(D1), _ptr(D2), _ptr(D3), _ptr(G), _ptr(loss)))

    def mlp_adam(self, W, G, M, V, step, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
        _check(_lib.oz2_mlp_adam(self._h, _ptr(W), _ptr(G), _ptr(M), _ptr(V), int(step), float(lr), float(beta1),
                                 float(beta2), float(eps), float(wd)))

    def batch_autotune(self, problems, bits=53, select=False):
        """-> dict(tiles, persistent, largest_first, ms): the fastest plan found by measuring candidates
        (overwrites the problems' C); pass the results on to Batch(...)"""
        n = len(problems)
        tiles = (_int * max(1, n))()
        fl = _int(0)
        ms = _dbl(0)
        _check(_lib.oz2_batch_autotune(self._h, n, _descs(problems), int(bits), BATCH_SELECT if select else 0, tiles,
                                       ctypes.byref(fl), ctypes.byref(ms)))
        return {"tiles": list(tiles)[:n], "persistent": bool(fl.value & BATCH_PERSISTENT),
                "largest_first": not (fl.value & BATCH_PROBLEM_ORDER), "ms": ms.value}
