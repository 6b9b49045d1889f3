"""ORACLE — test infrastructure only.

ctypes bindings to ``liboracle.so`` (the C restatement of the reference ALTO
path in ``alto_oracle.c``) and, when built, to ``_ref/libaltoref{64,128}.so``
(the reference's own ``include/alto.hpp`` / ``include/apr.hpp`` compiled from
its sources).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker or the timed CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so and, when the reference tree is present, _ref/:
    the reference's format + MTTKRP path, and its command line (src/main.cpp)
    built against the drop-in headers and linked with libalto_hip (when that
    library has been built)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/include")
    if with_ref:
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])
        if os.path.exists(os.path.join(os.path.dirname(_HERE), "alto_amd", "libalto_hip.so")):
            subprocess.check_call(["make", "-s", "-C", _HERE, "dropin"])


APR_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double)
APR_EV_START, APR_EV_PRE_FIX, APR_EV_POST_FIX, APR_EV_PHI, APR_EV_MU, APR_EV_NORM, APR_EV_OUTER = range(7)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build(with_ref=False)
        _LIB = load(path)
    return _LIB


def build_mutant(k: int, out: str) -> str:
    """An oracle build with planted mistake k (1: fmax in the KKT value,
    2: the kappa fix on Phi < 1, 3: 2-norm mode normalisation, 4: the
    outer stop test reversed), for tests that check the oracle's tests."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "mutant", f"MUTANT={int(k)}", f"MUTANT_OUT={out}"])
    return out


def load(path: str):
    """ctypes handle of an oracle build at `path`, with its signatures set."""
    if True:
        L = C.CDLL(path)
        L.or_pdep64.restype = C.c_uint64
        L.or_pdep64.argtypes = [C.c_uint64, C.c_uint64]
        L.or_pext64.restype = C.c_uint64
        L.or_pext64.argtypes = [C.c_uint64, C.c_uint64]
        L.or_pdep128.argtypes = [C.c_uint64, u64p, u64p]
        L.or_pext128.restype = C.c_uint64
        L.or_pext128.argtypes = [u64p, u64p]
        L.or_setup_masks.restype = C.c_int
        L.or_setup_masks.argtypes = [C.c_int, u64p, C.c_int, u64p, u64p]
        L.or_alto_create.restype = C.c_void_p
        L.or_alto_create.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int, C.c_int, C.c_int]
        L.or_alto_wrap.restype = C.c_void_p
        L.or_alto_wrap.argtypes = [C.c_int, u64p, C.c_uint64, C.c_int, C.c_int, u64p, u64p, f64p,
                                   u64p, u64p, u64p, C.c_void_p]
        L.or_alto_destroy.argtypes = [C.c_void_p]
        L.or_alto_get_masks.argtypes = [C.c_void_p, u64p, u64p, u64p, u64p]
        L.or_alto_get_idx.argtypes = [C.c_void_p, u64p]
        L.or_alto_get_vals.argtypes = [C.c_void_p, f64p]
        L.or_alto_get_prtn.argtypes = [C.c_void_p, u64p, u64p]
        L.or_alto_is_oidx.restype = C.c_int
        L.or_alto_is_oidx.argtypes = [C.c_void_p]
        L.or_alto_get_oidx.restype = C.c_int
        L.or_alto_get_oidx.argtypes = [C.c_void_p, C.c_int, u64p]
        L.or_alto_delinearize.argtypes = [C.c_void_p, u64p]
        L.or_mttkrp_alto.restype = C.c_int
        L.or_mttkrp_alto.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_uint64, C.c_int]
        L.or_mttkrp_coo.argtypes = [C.c_int, C.c_uint64, u64p, f64p, C.POINTER(C.c_void_p), C.c_uint64,
                                    C.c_int, f64p]
        L.or_random_init.argtypes = [C.c_int, u64p, C.c_uint64, C.c_uint, C.POINTER(C.c_void_p), f64p]
        L.or_gram.argtypes = [f64p, C.c_uint64, C.c_uint64, f64p]
        L.or_mat_norm.argtypes = [C.c_uint64, C.c_uint64, f64p, f64p, C.c_int]
        L.or_normalize_mode_l1.argtypes = [C.c_uint64, C.c_uint64, f64p, f64p]
        L.or_redistribute_lambda.argtypes = [C.c_uint64, C.c_uint64, f64p, f64p]
        L.or_pseudo_inverse.restype = C.c_int
        L.or_pseudo_inverse.argtypes = [C.c_int, C.c_uint64, C.POINTER(C.c_void_p), f64p, C.c_uint64, C.c_int]
        L.or_cpd_fit.restype = C.c_double
        L.or_cpd_fit.argtypes = [C.c_int, C.c_uint64, u64p, C.POINTER(C.c_void_p), f64p,
                                 C.POINTER(C.c_void_p), f64p, C.c_double]
        L.or_cpd_als.restype = C.c_int
        L.or_cpd_als.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_double, C.POINTER(C.c_void_p), f64p,
                                 f64p, C.c_int]
        L.or_phi.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_uint64, C.c_double, f64p]
        L.or_loglikelihood.restype = C.c_double
        L.or_loglikelihood.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), f64p, C.c_uint64, C.c_double]
        L.or_cp_apr.restype = C.c_int
        L.or_cp_apr.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.POINTER(C.c_void_p), f64p, C.c_void_p]
        L.or_cp_apr_hooked.restype = C.c_int
        L.or_cp_apr_hooked.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.POINTER(C.c_void_p), f64p, C.c_void_p,
                                       APR_HOOK, C.c_void_p]
        L.or_normsq.restype = C.c_double
        L.or_normsq.argtypes = [C.c_void_p]
        L.or_omp_threads.restype = C.c_int
        L.or_set_threads.argtypes = [C.c_int]
    return L


def ref_available(W: int = 1) -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", f"libaltoref{64 * W}.so"))


def ref(W: int = 1):
    if W not in _REF:
        L = C.CDLL(os.path.join(_HERE, "_ref", f"libaltoref{64 * W}.so"))
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int]
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_wrap.restype = C.c_void_p
        L.ref_wrap.argtypes = [C.c_int, u64p, C.c_uint64, u64p, u64p, u64p, f64p, C.c_int, u64p, u64p, u64p, u64p]
        L.ref_get.argtypes = [C.c_void_p, u64p, u64p, f64p, u64p, u64p, u64p, C.POINTER(C.c_int), u64p]
        L.ref_mttkrp_par.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_uint64]
        L.ref_mttkrp_seq.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_uint64]
        L.ref_phi_par.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_uint64, C.c_double, f64p]
        _REF[W] = L
    return _REF[W]


def ptrs(arrays):
    """array of void* to a list of C-contiguous numpy arrays"""
    return (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


class OracleAlto:
    """Owning wrapper of an oracle ALTO tensor handle."""

    def __init__(self, dims, cidx, vals, W=1, nprtn=8, val_int=False, handle=None, L=None):
        """L: an oracle build from ``load`` (default: liboracle.so)."""
        self.L = L if L is not None else lib()
        self.dims = np.ascontiguousarray(dims, dtype=np.uint64)
        self.nmode = len(self.dims)
        self.W = W
        self.nprtn = nprtn
        if handle is not None:
            self.h = handle
            self.nnz = int(len(vals))
            return
        cidx = np.ascontiguousarray(cidx, dtype=np.uint64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.nnz = int(vals.shape[0])
        self.h = self.L.or_alto_create(self.nmode, self.dims, self.nnz, cidx, vals, int(val_int), W, nprtn)
        if not self.h:
            raise ValueError("oracle create_alto failed")

    @classmethod
    def wrap(cls, dims, W, nprtn, masks, idx, vals, prtn_ptr, intervals, cr_masks, oidx=None, L=None):
        L = L if L is not None else lib()
        dims = np.ascontiguousarray(dims, dtype=np.uint64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        h = L.or_alto_wrap(len(dims), dims, len(vals), W, nprtn,
                               np.ascontiguousarray(masks, dtype=np.uint64),
                               np.ascontiguousarray(idx, dtype=np.uint64), vals,
                               np.ascontiguousarray(prtn_ptr, dtype=np.uint64),
                               np.ascontiguousarray(intervals, dtype=np.uint64),
                               np.ascontiguousarray(cr_masks, dtype=np.uint64),
                               None if oidx is None else np.ascontiguousarray(oidx, dtype=np.uint64).ctypes.data)
        return cls(dims, None, vals, W=W, nprtn=nprtn, handle=h, L=L)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.or_alto_destroy(h)
            self.h = None

    def state(self):
        N, W, P, nnz = self.nmode, self.W, self.nprtn, self.nnz
        masks = np.zeros(N * W, np.uint64)
        am = np.zeros(W, np.uint64)
        cr = np.zeros(N * W, np.uint64)
        acr = np.zeros(W, np.uint64)
        self.L.or_alto_get_masks(self.h, masks, am, cr, acr)
        idx = np.zeros(max(nnz, 1) * W, np.uint64)
        self.L.or_alto_get_idx(self.h, idx)
        vals = np.zeros(max(nnz, 1), np.float64)
        self.L.or_alto_get_vals(self.h, vals)
        pp = np.zeros(P + 1, np.uint64)
        iv = np.zeros(P * N * 2, np.uint64)
        self.L.or_alto_get_prtn(self.h, pp, iv)
        oidx = {}
        for n in range(N):
            o = np.zeros(max(nnz, 1), np.uint64)
            if self.L.or_alto_get_oidx(self.h, n, o):
                oidx[n] = o[:nnz]
        return dict(masks=masks.reshape(N, W), alto_mask=am, cr_masks=cr.reshape(N, W), alto_cr_mask=acr,
                    idx=idx[:nnz * W].reshape(nnz, W), vals=vals[:nnz], prtn_ptr=pp,
                    intervals=iv.reshape(P, N, 2), is_oidx=bool(self.L.or_alto_is_oidx(self.h)), oidx=oidx)

    def delinearize(self):
        c = np.zeros(max(self.nmode * self.nnz, 1), np.uint64)
        self.L.or_alto_delinearize(self.h, c)
        return c[: self.nmode * self.nnz].reshape(self.nmode, self.nnz)

    def mttkrp(self, mode, factors, variant=0, out=None):
        """Returns factors[mode] replaced by the mode-`mode` MTTKRP (zeroed first)."""
        R = factors[0].shape[1]
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        fs[mode] = np.zeros_like(fs[mode]) if out is None else out
        used = self.L.or_mttkrp_alto(self.h, mode, ptrs(fs), R, variant)
        if used < 0:
            raise ValueError("bad variant")
        return fs[mode]

    def phi(self, mode, factors, eps=1e-10):
        R = factors[0].shape[1]
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        Phi = np.zeros_like(fs[mode])
        self.L.or_phi(self.h, mode, ptrs(fs), R, eps, Phi)
        return Phi

    def normsq(self):
        return self.L.or_normsq(self.h)

    def cpd_als(self, factors, lam, max_iters, eps=1e-5, variant=0):
        R = factors[0].shape[1]
        fit = np.zeros(max_iters, np.float64)
        it = self.L.or_cpd_als(self.h, R, max_iters, eps, ptrs(factors), lam, fit, variant)
        return it, fit[:it]

    def cp_apr(self, factors, lam, max_iters, max_inner, eps=1e-5, eps_div_zero=1e-10, kappa_tol=1e-10,
               kappa=1e-2, dobench=False):
        R = factors[0].shape[1]
        buf = C.create_string_buffer(64)  # or_apr_result: 3 doubles, 7 ints
        it = self.L.or_cp_apr(self.h, R, max_iters, max_inner, eps, eps_div_zero, kappa_tol, kappa, int(dobench),
                             ptrs(factors), lam, C.cast(buf, C.c_void_p))
        d = np.frombuffer(buf.raw[:24], dtype=np.float64)
        ints = np.frombuffer(buf.raw[24:52], dtype=np.int32)
        return dict(iters=it, loglik=float(d[0]), ls_fit=float(d[1]), kkt=float(d[2]), outer_iters=int(ints[0]),
                    total_inner=int(ints[1]), inner_per_mode=[int(x) for x in ints[2:7]])

    def cp_apr_trace(self, factors, lam, max_iters, max_inner, eps=1e-5, eps_div_zero=1e-10, kappa_tol=1e-10,
                     kappa=1e-2, dobench=False):
        """cp_apr with the driver's steps recorded: returns (result, events),
        each event a dict(ev, iter, mode, inner, U (copies of every factor),
        lam, phi (copy or None), kkt)."""
        R = factors[0].shape[1]
        dims = [int(d) for d in self.dims]
        events = []

        def hook(ctx, ev, it, mode, inner, U, lamp, phi, kkt):
            Us = [np.ctypeslib.as_array(C.cast(U[m], C.POINTER(C.c_double)), shape=(dims[m], R)).copy()
                  for m in range(self.nmode)]
            lm = np.ctypeslib.as_array(lamp, shape=(R,)).copy()
            ph = None
            if phi and mode >= 0:
                ph = np.ctypeslib.as_array(phi, shape=(dims[mode], R)).copy()
            events.append(dict(ev=ev, iter=it, mode=mode, inner=inner, U=Us, lam=lm, phi=ph, kkt=kkt))

        cb = APR_HOOK(hook)
        buf = C.create_string_buffer(64)
        it = self.L.or_cp_apr_hooked(self.h, R, max_iters, max_inner, eps, eps_div_zero, kappa_tol, kappa,
                                     int(dobench), ptrs(factors), lam, C.cast(buf, C.c_void_p), cb, None)
        d = np.frombuffer(buf.raw[:24], dtype=np.float64)
        ints = np.frombuffer(buf.raw[24:52], dtype=np.int32)
        res = dict(iters=it, loglik=float(d[0]), ls_fit=float(d[1]), kkt=float(d[2]), outer_iters=int(ints[0]),
                   total_inner=int(ints[1]), inner_per_mode=[int(x) for x in ints[2:7]])
        return res, events


def random_init(dims, rank, seed):
    """KruskalModelRandomInit (glibc rand) — returns (factors, lambda)."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    U = [np.zeros((int(d), rank), np.float64) for d in dims]
    lam = np.zeros(rank, np.float64)
    lib().or_random_init(len(dims), dims, rank, seed, ptrs(U), lam)
    return U, lam


def setup_masks(dims, W):
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    m = np.zeros(len(dims) * W, np.uint64)
    am = np.zeros(W, np.uint64)
    bits = lib().or_setup_masks(len(dims), dims, W, m, am)
    if bits < 0:
        raise ValueError("mask setup failed")
    return m.reshape(len(dims), W), am, bits


def mttkrp_coo(dims, cidx, vals, factors, mode):
    cidx = np.ascontiguousarray(cidx, dtype=np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    out = np.zeros_like(fs[mode])
    lib().or_mttkrp_coo(len(dims), len(vals), cidx, vals, ptrs(fs), fs[0].shape[1], mode, out)
    return out


def cpd_coo(dims, cidx, vals, factors, lam, max_iters, eps=1e-5):
    """The reference's COO CP-ALS, ``cpd`` (include/cpd.hpp:274-371) with
    ``cpd_fit`` (:373-455), restated with numpy / scipy: an independent
    check of the ALTO-based ``OracleAlto.cpd_als`` (a different MTTKRP
    traversal and a LAPACK solve).  Per iteration and mode: U_j = MTTKRP_j
    V^-1 with V = prod of the other modes' Grams (Cholesky; least squares
    when it fails, include/cpd.hpp:546-646), columns normalised (2-norm in
    the first iteration, max(1, max_i U) after, src/kruskal_model.cpp:
    224-341), Gram updated; then the fit.  Factors and lambda are updated in
    place; returns (iterations, fits)."""
    import scipy.linalg

    N = len(dims)
    U = factors
    grams = [u.T @ u for u in U]
    normX = float(np.dot(vals, vals))
    fits, prev = [], 0.0
    scratch = None
    for it in range(max_iters):
        for j in range(N):
            Mj = mttkrp_coo(dims, cidx, vals, U, j)
            if j == N - 1:
                scratch = Mj.copy()
            V = np.ones_like(grams[0])
            for m in range(N):
                if m != j:
                    V *= grams[m]
            try:
                U[j][:] = scipy.linalg.cho_solve(scipy.linalg.cho_factor(V, lower=True), Mj.T).T
            except np.linalg.LinAlgError:
                U[j][:] = scipy.linalg.lstsq(V, Mj.T, lapack_driver="gelsy")[0].T
            if it == 0:
                lam[:] = np.sqrt((U[j] ** 2).sum(axis=0))
            else:
                lam[:] = np.maximum(1.0, np.maximum(U[j].max(axis=0), 0.0))
            U[j] /= lam
            grams[j] = U[j].T @ U[j]
        inner = float(((U[N - 1] * scratch).sum(axis=0) * lam).sum())
        tg = np.outer(lam, lam)
        for g in grams:
            tg = tg * g
        normU = abs(float(np.tril(tg, -1).sum() * 2 + np.trace(tg)))
        res = normX + normU - 2 * inner
        if res > 0.0:
            res = np.sqrt(res)
        fit = 1 - res / np.sqrt(normX)
        fits.append(fit)
        if it > 0 and abs(prev - fit) < eps:
            break
        prev = fit
    return len(fits), np.array(fits)


def cp_apr_coo(dims, cidx, vals, factors, lam, max_iters, max_inner, eps=1e-4, eps_div_zero=1e-10,
               kappa_tol=1e-10, kappa=1e-2):
    """The reference's COO CP-APR, ``cp_apr_mu_alto`` (src/apr.cpp:239-488),
    restated with numpy: the Khatri-Rao rows Pi are materialised per mode
    (``CalculatePi``, :213-235), Phi is a scatter-add over the coordinate
    list (``CalculatePhi``, :175-211), and the log-likelihood is the COO
    ``tt_logLikelihood`` (:95-173).  An independent traversal of the same
    algorithm as ``OracleAlto.cp_apr`` (which follows the ALTO driver,
    include/apr.hpp:984-1376); the two differ only in the outer stop: here
    the loop ends when every mode's first inner step met the KKT
    tolerance (:445-448).  Factors and lambda are updated in place.
    Returns dict(kkt (per outer iteration), inner (per outer iteration),
    outer_iters, total_inner, loglik, ls_fit)."""
    N = len(dims)
    cidx = np.asarray(cidx, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    U = factors
    R = U[0].shape[1]
    Phi = [np.zeros_like(u) for u in U]

    def l1_normalize(j):  # KruskalModelNormalize(Mode), src/kruskal_model.cpp:128-222
        cn = np.abs(U[j]).sum(axis=0)
        U[j] /= cn
        lam[:] = lam * cn

    for j in range(N):
        l1_normalize(j)
    kkts, inners = [], []
    it = 0
    for it in range(max_iters):
        converged = True
        n_in = 0
        kkt_mode = np.zeros(N)
        for n in range(N):
            B = U[n]
            if it > 0:  # src/apr.cpp:323-333
                fix = (B < kappa_tol) & (Phi[n] > 1.0)
                B[fix] = B[fix] + kappa
            B *= lam  # RedistributeLambda, src/kruskal_model.cpp:355-378
            lam[:] = 1.0
            Pi = np.ones((len(vals), R))
            for m in range(N):
                if m != n:
                    Pi *= U[m][cidx[m]]
            for _ in range(max_inner):
                n_in += 1
                den = np.maximum((B[cidx[n]] * Pi).sum(axis=1), eps_div_zero)
                Phi[n][:] = 0.0
                np.add.at(Phi[n], cidx[n], (vals / den)[:, None] * Pi)
                kv = float(np.abs(np.minimum(B, 1.0 - Phi[n])).max())
                kkt_mode[n] = kv
                if kv < eps:
                    break
                converged = False
                B *= Phi[n]
            l1_normalize(n)  # src/apr.cpp:387-398
        kkts.append(float(kkt_mode.max()))
        inners.append(n_in)
        if converged:
            break
    else:
        it = max_iters - 1  # src/apr.cpp:417-420
    total = int(sum(inners[: it + 1]))
    # tt_logLikelihood (COO), src/apr.cpp:95-173
    U[0] *= lam
    lam[:] = 1.0
    A = np.ones((len(vals), R))
    for m in range(N):
        A *= U[m][cidx[m]]
    t = np.maximum(A.sum(axis=1), eps_div_zero)
    loglik = float((vals * np.log(t)).sum() - U[0].sum())
    cn = np.abs(U[0]).sum(axis=0)
    U[0] /= cn
    lam *= cn
    # least-squares fit: Grams, last-mode MTTKRP, cpd_fit_ (src/apr.cpp:12-85)
    grams = [u.T @ u for u in U]
    Mlast = mttkrp_coo(dims, cidx.astype(np.uint64), vals, U, N - 1)
    inner_p = float(((U[N - 1] * Mlast).sum(axis=0) * lam).sum())
    tg = np.outer(lam, lam)
    for g in grams:
        tg = tg * g
    normU = abs(float(np.tril(tg, -1).sum() * 2 + np.trace(tg)))
    normX = float(np.dot(vals, vals))
    res = normX + normU - 2 * inner_p
    if res > 0.0:
        res = np.sqrt(res)
    return dict(kkt=np.array(kkts), inner=np.array(inners), outer_iters=it, total_inner=total, loglik=loglik,
                ls_fit=1 - res / np.sqrt(normX))


class RefAlto:
    """The reference's own create_alto / mttkrp_alto_par / CalculatePhi (oracle/_ref)."""

    def __init__(self, dims, cidx, vals, W=1, nprtn=8):
        self.L = ref(W)
        self.W = W
        self.dims = np.ascontiguousarray(dims, dtype=np.uint64)
        self.nmode = len(self.dims)
        self.nprtn = nprtn
        cidx = np.ascontiguousarray(cidx, dtype=np.uint64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.nnz = len(vals)
        self.h = self.L.ref_create(self.nmode, self.dims, self.nnz, cidx, vals, nprtn)

    @classmethod
    def wrap(cls, dims, W, nprtn, masks, alto_mask, idx, vals, prtn_ptr, intervals, cr_masks, alto_cr_mask):
        """The reference's AltoTensor over a given ALTO state (no create_alto)."""
        obj = cls.__new__(cls)
        obj.L = ref(W)
        obj.W = W
        obj.dims = np.ascontiguousarray(dims, dtype=np.uint64)
        obj.nmode = len(obj.dims)
        obj.nprtn = nprtn
        obj.nnz = len(vals)
        c = lambda a: np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
        obj.h = obj.L.ref_wrap(obj.nmode, obj.dims, obj.nnz, c(masks), c(alto_mask), c(idx),
                               np.ascontiguousarray(vals, dtype=np.float64), nprtn, c(prtn_ptr), c(intervals),
                               c(cr_masks), c(alto_cr_mask))
        return obj

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def state(self):
        N, W, P, nnz = self.nmode, self.W, self.nprtn, self.nnz
        masks = np.zeros(N * W, np.uint64)
        idx = np.zeros(nnz * W, np.uint64)
        vals = np.zeros(nnz, np.float64)
        pp = np.zeros(P + 1, np.uint64)
        iv = np.zeros(P * N * 2, np.uint64)
        cr = np.zeros(N * W, np.uint64)
        io = C.c_int(0)
        oidx = np.zeros(N * nnz, np.uint64)
        self.L.ref_get(self.h, masks, idx, vals, pp, iv, cr, C.byref(io), oidx)
        return dict(masks=masks.reshape(N, W), idx=idx.reshape(nnz, W), vals=vals, prtn_ptr=pp,
                    intervals=iv.reshape(P, N, 2), cr_masks=cr.reshape(N, W), is_oidx=bool(io.value),
                    oidx=oidx.reshape(N, nnz))

    def mttkrp(self, mode, factors, seq=False):
        fs = [np.ascontiguousarray(f, dtype=np.float64).copy() for f in factors]
        fs[mode][:] = 0
        (self.L.ref_mttkrp_seq if seq else self.L.ref_mttkrp_par)(self.h, mode, ptrs(fs), fs[0].shape[1])
        return fs[mode]

    def phi(self, mode, factors, eps=1e-10):
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        Phi = np.zeros_like(fs[mode])
        self.L.ref_phi_par(self.h, mode, ptrs(fs), fs[0].shape[1], eps, Phi)
        return Phi
