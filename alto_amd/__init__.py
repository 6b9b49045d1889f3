"""Python mirror of the ALTO API over libalto_hip (the MI355X C ABI).

Names follow the reference's interface (IntelLabs/ALTO): ``create_alto``
(include/alto.hpp:108), ``mttkrp_alto_par`` (include/alto.hpp:617),
``cpd_alto`` (include/cpd.hpp:43), ``cp_apr_alto`` (include/apr.hpp:984),
``KruskalModelRandomInit`` (src/kruskal_model.cpp:109).  All compute runs in
the HIP library; this module only marshals arguments.  Importing it without
the built library raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalto_hip.so")

ALTO_OK = 0
VAL_F64, VAL_U64 = 0, 1
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2
PTR_HOST, PTR_DEVICE = 0, 1
MAX_MODES = 5
APR_STOP_ALTO, APR_BENCH, APR_STOP_COO = 0, 1, 2
KRUSKAL_L1, KRUSKAL_NORM2, KRUSKAL_NORMMAX, KRUSKAL_REDISTRIBUTE = 0, 1, 2, 3

_lib = None


class AltoError(RuntimeError):
    pass


class TensorInfo(C.Structure):
    _fields_ = [("nmodes", C.c_int), ("lit_words", C.c_int), ("alto_bits", C.c_int), ("nprtn", C.c_int),
                ("is_oidx", C.c_int), ("val_kind", C.c_int), ("nnz", C.c_uint64),
                ("dims", C.c_uint64 * MAX_MODES), ("mode_masks", C.c_uint64 * (2 * MAX_MODES)),
                ("alto_mask", C.c_uint64 * 2), ("cr_masks", C.c_uint64 * (2 * MAX_MODES)),
                ("alto_cr_mask", C.c_uint64 * 2), ("device_bytes", C.c_uint64)]


class TuneEntry(C.Structure):
    _fields_ = [("mode", C.c_int), ("variant", C.c_int), ("seg", C.c_int), ("ms", C.c_double),
                ("default_ms", C.c_double)]


class CpdTimers(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("total", "mttkrp", "pseudoinv", "memcopy", "normalize", "update", "fit")]


class AprResult(C.Structure):
    _fields_ = [("loglik", C.c_double), ("ls_fit", C.c_double), ("kkt", C.c_double), ("outer_iters", C.c_int),
                ("total_inner", C.c_int), ("inner_per_mode", C.c_int * MAX_MODES), ("t_pre", C.c_double),
                ("t_apr", C.c_double), ("t_phi", C.c_double), ("t_post", C.c_double)]


def _one_hip_runtime():
    """Make this library and PyTorch share one HIP runtime, whichever of the
    two is imported first.

    PyTorch's ROCm wheel ships its own libamdhip64.so (soname
    libamdhip64.so.7), which its libraries request as "libamdhip64.so";
    libalto_hip.so requests "libamdhip64.so.7".  If this library were loaded
    first, the loader would map /opt/rocm's runtime for it and later torch's
    copy for torch (the names differ), and two HIP runtimes in one process
    do not share the GPU.  Loading torch's copy first, by path and with
    global symbols, makes both requests resolve to it: ours by soname, torch's
    by file identity.  Without PyTorch installed, /opt/rocm's runtime is
    used."""
    import importlib.util

    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)
    torch = sys.modules.get("torch")
    if torch is not None:  # already imported: let it initialise the runtime first
        try:
            torch.cuda.is_available()
        except Exception:
            pass


def lib():
    """Load libalto_hip.so; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AltoError(f"{LIB_PATH} is missing: build it with `python -m alto_amd.build`")
    _one_hip_runtime()
    L = C.CDLL(LIB_PATH)
    vp, u64, i32, f64 = C.c_void_p, C.c_uint64, C.c_int, C.c_double
    sig = {
        "alto_last_error": (C.c_char_p, []),
        "alto_version": (C.c_char_p, []),
        "alto_tensor_create": (i32, [i32, vp, u64, vp, vp, i32, i32, i32, i32, vp, C.POINTER(vp)]),
        "alto_tensor_destroy": (i32, [vp]),
        "alto_tensor_info": (i32, [vp, C.POINTER(TensorInfo)]),
        "alto_tensor_export": (i32, [vp, vp, vp, vp, vp]),
        "alto_tensor_delinearize": (i32, [vp, vp]),
        "alto_tensor_export_oidx": (i32, [vp, i32, vp]),
        "alto_pdep_keys": (i32, [vp, vp, u64, vp, vp]),
        "alto_pext_keys": (i32, [vp, vp, u64, vp, vp]),
        "alto_mttkrp": (i32, [vp, i32, u64, vp, vp, i32, i32, vp]),
        "alto_mttkrp_batched": (i32, [i32, vp, vp, u64, vp, vp, i32, i32, vp]),
        "alto_phi": (i32, [vp, i32, u64, vp, vp, f64, vp]),
        "alto_cpd_als": (i32, [vp, u64, i32, f64, i32, vp, vp, vp, C.POINTER(i32), C.POINTER(CpdTimers), i32, vp]),
        "alto_cpd_fit": (i32, [vp, u64, vp, vp, C.POINTER(f64)]),
        "alto_cp_apr": (i32, [vp, u64, i32, i32, f64, f64, f64, f64, i32, vp, vp, C.POINTER(AprResult), i32, vp]),
        "alto_cp_apr_ex": (i32, [vp, u64, i32, i32, f64, f64, f64, f64, i32, vp, vp, C.POINTER(AprResult), i32,
                                 vp]),
        "alto_cpd_fit_parts": (i32, [i32, vp, u64, vp, vp, vp, vp, f64, i32, C.POINTER(f64)]),
        "alto_kruskal_normalize": (i32, [i32, vp, u64, vp, vp, i32, i32, i32, vp]),
        "alto_coo_mttkrp": (i32, [i32, vp, u64, vp, vp, i32, u64, vp, vp, i32, i32, vp]),
        "alto_update_gram": (i32, [vp, u64, u64, i32, vp, vp]),
        "alto_update_gram_noalloc": (i32, [vp, u64, u64, i32, vp, vp]),
        "alto_cpd_session_create": (i32, [vp, u64, i32, vp, vp, C.POINTER(vp)]),
        "alto_cpd_session_create_seeded": (i32, [vp, u64, i32, u64, vp, C.POINTER(vp)]),
        "alto_cpd_session_iterate": (i32, [vp, i32, f64, vp, C.POINTER(i32), C.POINTER(i32), i32]),
        "alto_cpd_session_get": (i32, [vp, vp, vp, C.POINTER(f64), C.POINTER(CpdTimers)]),
        "alto_cpd_sessions_iterate": (i32, [i32, vp, i32, f64, vp, vp]),
        "alto_cpd_session_destroy": (i32, [vp]),
        "alto_pseudo_inverse": (i32, [vp, i32, i32, u64, vp, u64, C.POINTER(i32), vp]),
        "alto_tune_set": (i32, [i32, i32]),
        "alto_mttkrp_autotune": (i32, [vp, u64, i32, vp, vp, i32, vp, vp]),
        "alto_kruskal_random_init": (i32, [i32, vp, u64, C.c_uint, vp, vp]),
        "alto_device_malloc": (i32, [C.POINTER(vp), u64]),
        "alto_device_free": (i32, [vp]),
        "alto_memcpy_h2d": (i32, [vp, vp, u64]),
        "alto_memcpy_d2h": (i32, [vp, vp, u64]),
        "alto_memcpy_d2d": (i32, [vp, vp, u64, vp]),
        "alto_device_synchronize": (i32, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status != ALTO_OK:
        msg = lib().alto_last_error()
        raise AltoError(f"{what} failed ({status}): {msg.decode() if msg else ''}")


def _ptr_array(ptrs):
    return (C.c_void_p * len(ptrs))(*ptrs)


class DeviceArray:
    """Owning device buffer managed through the library's allocator."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().alto_device_malloc(C.byref(p), max(self.nbytes, 8)), "device malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().alto_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes), "h2d")
        return d

    def numpy(self):
        a = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().alto_memcpy_d2h(a.ctypes.data, self.ptr, self.nbytes), "d2h")
        return a

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().alto_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "h2d")

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _lib is not None:
            _lib.alto_device_free(p)
            self.ptr = None


def synchronize():
    check(lib().alto_device_synchronize(), "synchronize")


def to_bf16_bits(a):
    """float64/32 array -> uint16 bf16 bit patterns (round to nearest even)."""
    f = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


class AltoTensor:
    """create_alto<LIT> (include/alto.hpp:108-294) on the device."""

    def __init__(self, dims, cidx, vals, nprtn=8, lit_bits=0, val_kind=VAL_F64, device_ptrs=False):
        self.dims = np.ascontiguousarray(dims, dtype=np.uint64)
        self.nmodes = len(self.dims)
        self.h = None
        if device_ptrs:
            rows = list(cidx)
            vp = vals
            kind = PTR_DEVICE
            nnz = int(self._nnz_hint)
        else:
            cidx = np.ascontiguousarray(cidx, dtype=np.uint64)
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            nnz = int(vals.shape[0])
            self._keep = (cidx, vals)
            rows = [cidx[m].ctypes.data for m in range(self.nmodes)]
            vp = vals.ctypes.data
            kind = PTR_HOST
        h = C.c_void_p()
        check(lib().alto_tensor_create(self.nmodes, self.dims.ctypes.data, nnz, _ptr_array(rows), vp, kind, val_kind,
                                       lit_bits, nprtn, None, C.byref(h)), "alto_tensor_create")
        self.h = h.value
        self._keep = None
        self.nnz = nnz
        self.nprtn = nprtn
        inf = self.info()
        self.W = inf.lit_words

    @classmethod
    def from_device(cls, dims, cidx_ptrs, vals_ptr, nnz, nprtn=8, lit_bits=0, val_kind=VAL_F64):
        obj = cls.__new__(cls)
        obj._nnz_hint = nnz
        cls.__init__(obj, dims, cidx_ptrs, vals_ptr, nprtn=nprtn, lit_bits=lit_bits, val_kind=val_kind,
                     device_ptrs=True)
        return obj

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.alto_tensor_destroy(h)
            self.h = None

    def info(self) -> TensorInfo:
        inf = TensorInfo()
        check(lib().alto_tensor_info(self.h, C.byref(inf)), "alto_tensor_info")
        return inf

    def export(self):
        N, W, P, nnz = self.nmodes, self.W, self.nprtn, self.nnz
        idx = np.zeros(nnz * W, np.uint64)
        vals = np.zeros(nnz, np.float64)
        pp = np.zeros(P + 1, np.uint64)
        iv = np.zeros(P * N * 2, np.uint64)
        check(lib().alto_tensor_export(self.h, idx.ctypes.data, vals.ctypes.data, pp.ctypes.data, iv.ctypes.data),
              "alto_tensor_export")
        inf = self.info()
        masks = np.array(inf.mode_masks[: N * W], np.uint64).reshape(N, W)
        cr = np.array(inf.cr_masks[: N * W], np.uint64).reshape(N, W)
        return dict(masks=masks, idx=idx.reshape(nnz, W), vals=vals, prtn_ptr=pp, intervals=iv.reshape(P, N, 2),
                    cr_masks=cr, is_oidx=bool(inf.is_oidx), alto_bits=inf.alto_bits)

    def delinearize(self):
        c = np.zeros(self.nmodes * self.nnz, np.uint64)
        check(lib().alto_tensor_delinearize(self.h, c.ctypes.data), "delinearize")
        return c.reshape(self.nmodes, self.nnz)

    def oidx(self, mode):
        o = np.zeros(self.nnz, np.uint64)
        check(lib().alto_tensor_export_oidx(self.h, mode, o.ctypes.data), "export_oidx")
        return o

    # ---- MTTKRP ----
    def mttkrp_device(self, mode, factors, out, precision=PREC_F64, accumulate=False, stream=None):
        ptrs = _ptr_array([f.ptr for f in factors])
        check(lib().alto_mttkrp(self.h, mode, factors[0].shape[1], ptrs, out.ptr, precision, int(accumulate), stream),
              "alto_mttkrp")

    def mttkrp_autotune(self, rank, precision=PREC_F64, reps=3, factor_ptrs=None, out_ptr=None, stream=None):
        """alto_mttkrp_autotune: picks the MTTKRP launch shape per mode for this
        (rank, precision) and keeps it on the tensor.  Without device factors
        given, random ones of that precision are used.  Returns one dict per mode."""
        keep = []
        if factor_ptrs is None:
            rng = np.random.default_rng(0)
            for d in self.dims:
                u = rng.random((int(d), rank))
                a = (DeviceArray.from_numpy(u) if precision == PREC_F64 else
                     DeviceArray.from_numpy(u.astype(np.float32)) if precision == PREC_F32 else
                     DeviceArray.from_numpy(to_bf16_bits(u)))
                keep.append(a)
            factor_ptrs = [a.ptr for a in keep]
        if out_ptr is None:
            o = DeviceArray((int(max(self.dims)), rank), np.float64)
            keep.append(o)
            out_ptr = o.ptr
        res = (TuneEntry * self.nmodes)()
        check(lib().alto_mttkrp_autotune(self.h, rank, precision, _ptr_array(list(factor_ptrs)), out_ptr, reps, stream,
                                         C.cast(res, C.c_void_p)), "alto_mttkrp_autotune")
        return [{k: getattr(r, k) for k, _ in TuneEntry._fields_} for r in res]

    def mttkrp(self, mode, factors, precision=PREC_F64):
        """mttkrp_alto_par on host arrays: returns the mode-`mode` MTTKRP."""
        R = factors[0].shape[1]
        if precision == PREC_F64:
            dev = [DeviceArray.from_numpy(np.asarray(f, np.float64)) for f in factors]
            out = DeviceArray((int(self.dims[mode]), R), np.float64)
        elif precision == PREC_F32:
            dev = [DeviceArray.from_numpy(np.asarray(f, np.float32)) for f in factors]
            out = DeviceArray((int(self.dims[mode]), R), np.float64)
        else:
            dev = [DeviceArray.from_numpy(to_bf16_bits(f)) for f in factors]
            out = DeviceArray((int(self.dims[mode]), R), np.float64)
        self.mttkrp_device(mode, dev, out, precision)
        synchronize()
        return out.numpy()

    def phi(self, mode, factors, eps=1e-10):
        R = factors[0].shape[1]
        dev = [DeviceArray.from_numpy(np.asarray(f, np.float64)) for f in factors]
        out = DeviceArray((int(self.dims[mode]), R), np.float64)
        check(lib().alto_phi(self.h, mode, R, _ptr_array([d.ptr for d in dev]), out.ptr, eps, None), "alto_phi")
        synchronize()
        return out.numpy()

    # ---- CP-ALS / CP-APR ----
    def cpd_als(self, factors, lam, max_iters, epsilon=1e-5, precision=PREC_F64, verbose=False):
        """cpd_alto: factors (list of float64 arrays) and lam are updated in place."""
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        for i, f in enumerate(fs):
            if f is not factors[i]:
                raise ValueError("factors must be C-contiguous float64 arrays")
        fit = np.zeros(max_iters, np.float64)
        iters = C.c_int(0)
        tm = CpdTimers()
        check(lib().alto_cpd_als(self.h, fs[0].shape[1], max_iters, epsilon, precision,
                                 _ptr_array([f.ctypes.data for f in fs]), lam.ctypes.data, fit.ctypes.data,
                                 C.byref(iters), C.byref(tm), int(verbose), None), "alto_cpd_als")
        return iters.value, fit[: iters.value], {k: getattr(tm, k) for k, _ in CpdTimers._fields_}

    def cp_apr(self, factors, lam, max_iters, max_inner, epsilon=1e-5, eps_div_zero=1e-10, kappa_tol=1e-10,
               kappa=1e-2, dobench=False, verbose=False, stop_rule=None):
        """cp_apr_alto (stop_rule None: APR_BENCH when dobench, else
        APR_STOP_ALTO) or the COO drivers' rule (APR_STOP_COO)."""
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        for i, f in enumerate(fs):
            if f is not factors[i]:
                raise ValueError("factors must be C-contiguous float64 arrays")
        res = AprResult()
        if stop_rule is None:
            stop_rule = APR_BENCH if dobench else APR_STOP_ALTO
        check(lib().alto_cp_apr_ex(self.h, fs[0].shape[1], max_iters, max_inner, epsilon, eps_div_zero, kappa_tol,
                                   kappa, int(stop_rule), _ptr_array([f.ctypes.data for f in fs]), lam.ctypes.data,
                                   C.byref(res), int(verbose), None), "alto_cp_apr_ex")
        return dict(loglik=res.loglik, ls_fit=res.ls_fit, kkt=res.kkt, outer_iters=res.outer_iters,
                    total_inner=res.total_inner, inner_per_mode=list(res.inner_per_mode),
                    t_pre=res.t_pre, t_apr=res.t_apr, t_phi=res.t_phi, t_post=res.t_post)


class CpdSession:
    """Device-resident CP-ALS model over an AltoTensor (alto_cpd_session_*)."""

    def __init__(self, tensor, factors, precision=PREC_F64, stream=None):
        self.t = tensor
        self.R = factors[0].shape[1]
        self.dims = [f.shape[0] for f in factors]
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        h = C.c_void_p()
        check(lib().alto_cpd_session_create(tensor.h, self.R, precision, _ptr_array([f.ctypes.data for f in fs]),
                                            stream, C.byref(h)), "cpd_session_create")
        self.h = h.value

    def iterate(self, n, epsilon=0.0, verbose=False):
        fit = np.zeros(max(n, 1), np.float64)
        ran, conv = C.c_int(0), C.c_int(0)
        check(lib().alto_cpd_session_iterate(self.h, n, epsilon, fit.ctypes.data, C.byref(ran), C.byref(conv),
                                             int(verbose)), "cpd_session_iterate")
        return fit[: ran.value], bool(conv.value)

    def get(self):
        U = [np.zeros((d, self.R), np.float64) for d in self.dims]
        lam = np.zeros(self.R, np.float64)
        fit = C.c_double(0)
        tm = CpdTimers()
        check(lib().alto_cpd_session_get(self.h, _ptr_array([u.ctypes.data for u in U]), lam.ctypes.data,
                                         C.byref(fit), C.byref(tm)), "cpd_session_get")
        return U, lam, fit.value, {k: getattr(tm, k) for k, _ in CpdTimers._fields_}

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.alto_cpd_session_destroy(h)
            self.h = None


def coo_mttkrp(dims, cidx, vals, mode, factors, out=None):
    """alto_coo_mttkrp on host arrays: out (dims[mode] x R) += the mode-`mode`
    MTTKRP of the coordinate tensor (cidx: nmodes x nnz uint64); a new zero
    output when out is None.  Returns out."""
    cidx = np.ascontiguousarray(cidx, np.uint64)
    vals = np.ascontiguousarray(vals, np.float64)
    fs = [np.ascontiguousarray(f, np.float64) for f in factors]
    R = fs[0].shape[1]
    acc = out is not None
    if out is None:
        out = np.zeros((dims[mode], R), np.float64)
    if not (out.flags.c_contiguous and out.dtype == np.float64):
        raise ValueError("out must be a C-contiguous float64 array")
    d = (C.c_uint64 * len(dims))(*dims)
    check(lib().alto_coo_mttkrp(len(dims), d, vals.size, _ptr_array([cidx[m].ctypes.data for m in range(len(dims))]),
                                vals.ctypes.data, mode, R, _ptr_array([f.ctypes.data for f in fs]),
                                out.ctypes.data, int(acc), PTR_HOST, None), "alto_coo_mttkrp")
    return out


def kruskal_normalize(factors, lam, mode, kind):
    """alto_kruskal_normalize on host arrays (updated in place): KRUSKAL_L1
    (mode -1: every mode), KRUSKAL_NORM2, KRUSKAL_NORMMAX, KRUSKAL_REDISTRIBUTE."""
    for f in list(factors) + [lam]:
        if not (f.flags.c_contiguous and f.dtype == np.float64):
            raise ValueError("factors and lambda must be C-contiguous float64 arrays")
    dims = (C.c_uint64 * len(factors))(*[f.shape[0] for f in factors])
    check(lib().alto_kruskal_normalize(len(factors), dims, factors[0].shape[1],
                                       _ptr_array([f.ctypes.data for f in factors]), lam.ctypes.data, mode, kind,
                                       PTR_HOST, None), "alto_kruskal_normalize")


def cpd_fit_parts(grams, lam, U_last, U_mttkrp, normX, dims):
    """alto_cpd_fit_parts on host arrays: the fit of cpd_fit_alto from its
    arguments (Grams, lambda, last factor, last-mode MTTKRP, ||X||^2)."""
    gs = [np.ascontiguousarray(g, np.float64) for g in grams]
    lam = np.ascontiguousarray(lam, np.float64)
    U_last = np.ascontiguousarray(U_last, np.float64)
    U_mttkrp = np.ascontiguousarray(U_mttkrp, np.float64)
    d = (C.c_uint64 * len(dims))(*dims)
    fit = C.c_double(0)
    check(lib().alto_cpd_fit_parts(len(dims), d, lam.size, _ptr_array([g.ctypes.data for g in gs]), lam.ctypes.data,
                                   U_last.ctypes.data, U_mttkrp.ctypes.data, float(normX), PTR_HOST,
                                   C.byref(fit)), "alto_cpd_fit_parts")
    return fit.value


def cpd_sessions_iterate(sessions, n, epsilon=0.0, stream=None):
    """alto_cpd_sessions_iterate: the sessions (one rank and precision) run up
    to n CP-ALS iterations in lockstep; returns the iterations run per session."""
    k = len(sessions)
    ran = (C.c_int * max(k, 1))()
    check(lib().alto_cpd_sessions_iterate(k, _ptr_array([s.h for s in sessions]), n, epsilon, stream,
                                          C.cast(ran, C.c_void_p)), "cpd_sessions_iterate")
    return [ran[i] for i in range(k)]


def _factor_arrays(factors, precision):
    if precision == PREC_F64:
        return [DeviceArray.from_numpy(np.asarray(f, np.float64)) for f in factors]
    if precision == PREC_F32:
        return [DeviceArray.from_numpy(np.asarray(f, np.float32)) for f in factors]
    return [DeviceArray.from_numpy(to_bf16_bits(f)) for f in factors]


def mttkrp_batched(tensors, modes, factors, precision=PREC_F64):
    """alto_mttkrp_batched on host arrays: job i is the mode-modes[i] MTTKRP
    of tensors[i] (an AltoTensor) with factors[i] (its host factor list); all
    jobs share the rank.  Returns the list of outputs."""
    n = len(tensors)
    if not (len(modes) == len(factors) == n):
        raise ValueError("tensors, modes and factors differ in length")
    R = factors[0][0].shape[1] if n else 1
    keep, fptrs, outs = [], [], []
    for t, m, fs in zip(tensors, modes, factors):
        dev = _factor_arrays(fs, precision)
        arr = _ptr_array([d.ptr for d in dev])
        keep.append((dev, arr))
        fptrs.append(C.cast(arr, C.c_void_p).value)
        rows = int(t.dims[m]) if 0 <= m < len(t.dims) else 1  # a bad mode is the library's to reject
        outs.append(DeviceArray((rows, R), np.float64))
    th = _ptr_array([t.h for t in tensors])
    md = (C.c_int * n)(*modes)
    check(lib().alto_mttkrp_batched(n, th, md, R, _ptr_array(fptrs), _ptr_array([o.ptr for o in outs]), precision,
                                    0, None), "alto_mttkrp_batched")
    synchronize()
    return [o.numpy() for o in outs]


def create_alto(dims, cidx, vals, nprtn=8, lit_bits=0, val_kind=VAL_F64):
    return AltoTensor(dims, cidx, vals, nprtn=nprtn, lit_bits=lit_bits, val_kind=val_kind)


def kruskal_random_init(dims, rank, seed):
    """KruskalModelRandomInit (src/kruskal_model.cpp:109-125): glibc srand/rand."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    U = [np.zeros((int(d), rank), np.float64) for d in dims]
    lam = np.zeros(rank, np.float64)
    check(lib().alto_kruskal_random_init(len(dims), dims.ctypes.data, rank, seed,
                                         _ptr_array([u.ctypes.data for u in U]), lam.ctypes.data), "random_init")
    return U, lam
