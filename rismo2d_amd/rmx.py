"""ctypes binding of librmx.so (the C ABI declared in include/rmx.h).

This is the only way Python reaches the GPU path: there is no fallback.  Loading fails
with an ImportError-like ``RmxError`` if librmx.so has not been built, and every call
that fails in the library raises ``RmxError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "rmx.h")

F64, F32 = 64, 32
BICGSTAB, FGMRES = 1, 2
PREC_NONE, PREC_ILU0, PREC_MPNN, PREC_SILU0, PREC_SILU0C = 0, 1, 2, 3, 4
ERR_NODEVICE = -4

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int)
_v = C.c_void_p


class RmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rmx error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("np", C.c_int), ("x", _d), ("y", _d), ("z", _d), ("ne", C.c_int),
                ("elems", _i), ("nnd", _i), ("manning", _d), ("vt", _d),
                ("n_inlet", C.c_int), ("inlet_segs", _i), ("n_outlet", C.c_int),
                ("outlet_segs", _i), ("Q", C.c_double), ("S_in", C.c_double),
                ("S_out", C.c_double), ("veg_dp", _d), ("veg_sp", _d), ("turb_st", _d),
                ("turb_kdest", _d), ("turb_estx", _d), ("turb_esty", _d)]


class Params(C.Structure):
    _fields_ = [("relax_method", C.c_int), ("dry_method", C.c_int), ("dry_limit", C.c_double),
                ("max_delta_uv", C.c_double), ("max_delta_s", C.c_double),
                ("conv_uv", C.c_double), ("conv_s", C.c_double), ("max_us", C.c_double),
                ("turb", C.c_int), ("cm", C.c_double), ("cd", C.c_double),
                ("relax_time", C.c_double), ("interval", C.c_double),
                ("rewet_limit", C.c_double), ("count_down", C.c_int)]


class KrylovOpts(C.Structure):
    pass


KrylovOpts._fields_ = [("method", C.c_int), ("prec", C.c_int), ("precond", C.c_int),
                       ("max_iter", C.c_int), ("mkyrl", C.c_int), ("eps", C.c_double),
                       ("refine", C.c_int), ("refine_tol", C.c_double), ("refine_max", C.c_int),
                       ("proceed", C.c_int), ("next", C.POINTER(KrylovOpts))]


def krylov_opts(method=1, prec=64, precond=1, max_iter=1000, mkyrl=30, eps=1e-10, refine=False,
                refine_tol=1e-8, refine_max=10, proceed=0, next=None):
    """A KrylovOpts chain: ``next`` (a dict of the same keywords) is the solver that follows
    an unconverged solve when ``proceed`` > 0 (the reference's SOLVER::proceed).  The
    returned structure keeps its successors alive."""
    o = KrylovOpts(method, prec, precond, max_iter, mkyrl, eps, int(refine), refine_tol,
                   refine_max, proceed, None)
    if next is not None:
        o._next = krylov_opts(**next)
        o.next = C.pointer(o._next)
    return o


class KrylovOut(C.Structure):
    _fields_ = [("iters", C.c_int), ("sweeps", C.c_int), ("converged", C.c_int),
                ("accuracy", C.c_double), ("true_resid", C.c_double), ("seconds", C.c_double),
                ("fallback", C.c_int), ("solvers", C.c_int), ("interrupted", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("maxAbs", C.c_double * 3), ("avAbs", C.c_double * 3), ("relaxUs", C.c_double),
                ("relaxS", C.c_double), ("nrconv", C.c_int), ("scale", C.c_double),
                ("solve", KrylovOut), ("t_prepare", C.c_double), ("t_assemble", C.c_double),
                ("t_factor", C.c_double), ("t_solve", C.c_double), ("t_update", C.c_double),
                ("relax_time", C.c_double), ("relax_time_last", C.c_double)]


class SweepInfo(C.Structure):
    _fields_ = [("lower_entries", C.c_longlong), ("upper_entries", C.c_longlong),
                ("lower_levels", C.c_int), ("upper_levels", C.c_int), ("strips", C.c_int),
                ("segments", C.c_int)]


class KernelTimes(C.Structure):
    _fields_ = [("spmv_s", C.c_double), ("ilu_apply_s", C.c_double), ("ilu_fwd_s", C.c_double),
                ("ilu_bwd_s", C.c_double), ("silu_fwd_s", C.c_double), ("silu_bwd_s", C.c_double),
                ("n_spmv", C.c_longlong), ("n_ilu", C.c_longlong), ("spmv_bytes", C.c_double)]


class KernelStats(C.Structure):
    _fields_ = [("seconds", C.c_double * 8), ("bytes", C.c_double * 8),
                ("launches", C.c_longlong * 8)]


# kernel classes of rmx_kernel_stats
KERNEL_CLASSES = ("ilu_factor", "sweep_lower", "sweep_upper", "spmv", "coarse", "gram_schmidt",
                  "resid_fp64", "sweeps_block")

_LIB = None


def header_functions():
    """Names of the entry points declared in include/rmx.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:rmx_status|const char\*)\s+(rmx_\w+)\s*\(", txt, re.M)))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RmxError(-100, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.rmx_last_error.restype = C.c_char_p
        sig = {
            "rmx_create": [C.c_int, C.POINTER(_v)], "rmx_destroy": [_v],
            "rmx_mesh_upload": [_v, C.POINTER(MeshDesc)], "rmx_set_params": [_v, C.POINTER(Params)],
            "rmx_state_upload": [_v, _d, _d, _d], "rmx_state_download": [_v, _d, _d, _d, _d],
            "rmx_prepare": [_v, _i, C.POINTER(C.c_longlong)], "rmx_get_numbering": [_v, _i],
            "rmx_get_pattern": [_v, _i, _i], "rmx_get_nodal": [_v, _d, _d, _i, _d],
            "rmx_n_model_elems": [_v, _i], "rmx_element_coefs": [_v, C.c_int, C.c_int, _d, _d],
            "rmx_turbulence": [_v, _d, _d],
            "rmx_time_step": [_v, C.c_double, C.c_double, C.c_int], "rmx_time_commit": [_v],
            "rmx_slip_flow": [_v], "rmx_dry_rewet": [_v, _i, _i],
            "rmx_get_wetdry": [_v, _i, _i], "rmx_get_relax_time": [_v, _d, _d, _d],
            "rmx_time_state": [_v, C.c_int, C.POINTER(_d)],
            "rmx_set_inflow": [_v, C.c_double, C.c_double, C.c_int],
            "rmx_hydrograph": [C.c_int, _d, _d, _d, C.c_double, _d, _d],
            "rmx_assemble": [_v, C.c_int], "rmx_scale_l2": [_v, _d],
            "rmx_convert_matrix": [_v, C.c_int], "rmx_ilu0_factor": [_v, C.c_int],
            "rmx_ilu0_apply": [_v, C.c_int, _d, _d], "rmx_spmv": [_v, C.c_int, _d, _d],
            "rmx_silu0_factor": [_v, C.c_int], "rmx_silu0_apply": [_v, C.c_int, _d, _d],
            "rmx_set_strips": [_v, C.c_int, C.c_int], "rmx_get_blocks": [_v, _i],
            "rmx_mpnn_n_params": [_i], "rmx_mpnn_set_weights": [_v, _f, C.c_int],
            "rmx_mpnn_get_weights": [_v, _f, C.c_int], "rmx_mpnn_forward": [_v, _d, _d, _f],
            "rmx_mpnn_add_sample": [_v, _d, C.c_int, _i],
            "rmx_mpnn_loss_grad": [_v, C.c_int, _d, _f],
            "rmx_mpnn_train": [_v, C.c_int, C.c_double, _d],
            "rmx_set_coarse": [_v, C.c_int], "rmx_coarse_factor": [_v, C.c_int],
            "rmx_get_aggregates": [_v, _i, _i], "rmx_precond_apply": [_v, C.c_int, C.c_int, _d, _d],
            "rmx_get_sweep_info": [_v, C.c_int, C.POINTER(SweepInfo)],
            "rmx_solve": [_v, C.POINTER(KrylovOpts), C.POINTER(KrylovOut)],
            "rmx_get_matrix": [_v, C.c_int, _v, _d], "rmx_get_factors": [_v, C.c_int, _v],
            "rmx_get_solution": [_v, _d], "rmx_set_rhs": [_v, _d], "rmx_set_solution": [_v, _d],
            "rmx_csr_upload": [_v, C.c_int, _i, _i, _d, _d],
            "rmx_newton_update": [_v, C.POINTER(StepStats)],
            "rmx_newton_step": [_v, C.POINTER(KrylovOpts), C.POINTER(StepStats)],
            "rmx_profile_kernels": [_v, C.c_int, C.c_int, C.POINTER(KernelTimes)],
            "rmx_kernel_stats_enable": [_v, C.c_int],
            "rmx_kernel_stats_get": [_v, C.POINTER(KernelStats)],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _LIB = L
    return _LIB


def _check(code):
    if code != 0:
        raise RmxError(code, lib().rmx_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Context:
    """One equation system on one GPU (an rmx_ctx)."""

    def __init__(self, device: int = 0):
        self.h = _v()
        _check(lib().rmx_create(device, C.byref(self.h)))
        self.np = 0
        self.ne = 0
        self.n = 0
        self.nnz = 0
        self._keep = {}

    def close(self):
        if self.h:
            lib().rmx_destroy(self.h)
            self.h = _v()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- model group -------------------------------------------------------------------
    def upload_mesh(self, deck):
        k = dict(x=_c(deck.x, np.float64), y=_c(deck.y, np.float64), z=_c(deck.z, np.float64),
                 el=_c(deck.elems, np.int32), nnd=_c(deck.nnd, np.int32),
                 man=_c(deck.manning, np.float64), vt=_c(deck.vt, np.float64),
                 inl=_c(deck.lines[1], np.int32), out=_c(deck.lines[2], np.int32))
        self._keep = k
        d = MeshDesc(deck.n_nodes, _p(k["x"], _d), _p(k["y"], _d), _p(k["z"], _d), deck.n_elems,
                     _p(k["el"], _i), _p(k["nnd"], _i), _p(k["man"], _d), _p(k["vt"], _d),
                     len(k["inl"]), _p(k["inl"], _i), len(k["out"]), _p(k["out"], _i),
                     deck.Q_inlet, deck.S_inlet, deck.S_outlet)
        if getattr(deck, "veg_dp", None) is not None:
            k["vdp"] = _c(deck.veg_dp, np.float64)
            k["vsp"] = _c(deck.veg_sp, np.float64)
            d.veg_dp = _p(k["vdp"], _d)
            d.veg_sp = _p(k["vsp"], _d)
        if getattr(deck, "turb_st", None) is not None:
            k["tst"] = _c(deck.turb_st, np.float64)
            d.turb_st = _p(k["tst"], _d)
        if getattr(deck, "turb_kdest", None) is not None:
            k["tkd"] = _c(deck.turb_kdest, np.float64)
            d.turb_kdest = _p(k["tkd"], _d)
        for key, name in (("tex", "turb_estx"), ("tey", "turb_esty")):
            if getattr(deck, name, None) is not None:
                k[key] = _c(getattr(deck, name), np.float64)
                setattr(d, name, _p(k[key], _d))
        _check(lib().rmx_mesh_upload(self.h, C.byref(d)))
        self.np = deck.n_nodes
        self.ne = deck.n_elems
        # the deck's turbulence model becomes active with the other parameters kept
        self.deck_turb = int(getattr(deck, "turb", 1))
        kw = dict(getattr(self, "_params", {}))
        kw["turb"] = self.deck_turb
        self.set_params(**kw)
        self.state_upload(deck.U0, deck.V0, deck.S0)

    def set_params(self, relax_method=2, dry_method=2, dry_limit=1e-3, max_delta_uv=10.0,
                   max_delta_s=10.0, conv_uv=1e-7, conv_s=1e-7, max_us=100.0, turb=None,
                   cm=0.09, cd=1.0, relax_time=0.0, interval=0.0, rewet_limit=0.005,
                   count_down=5):
        """turb: $TM_TURBULENCE code; None takes the uploaded deck's.  relax_time
        (relaxTimeFlow) and interval (the stationary cycle's time step): relaxation methods
        3 and 4.  rewet_limit, count_down: drying and rewetting ($DRYREW)."""
        if turb is None:
            turb = getattr(self, "deck_turb", 1)
        p = Params(relax_method, dry_method, dry_limit, max_delta_uv, max_delta_s, conv_uv,
                   conv_s, max_us, turb, cm, cd, relax_time, interval, rewet_limit,
                   count_down)
        _check(lib().rmx_set_params(self.h, C.byref(p)))
        # kept (and re-applied by upload_mesh) only once the library has accepted them
        self._params = dict(relax_method=relax_method, dry_method=dry_method,
                            dry_limit=dry_limit, max_delta_uv=max_delta_uv,
                            max_delta_s=max_delta_s, conv_uv=conv_uv, conv_s=conv_s,
                            max_us=max_us, cm=cm, cd=cd, relax_time=relax_time,
                            interval=interval, rewet_limit=rewet_limit, count_down=count_down)

    TIME_KEYS = ("Uo", "Vo", "So", "dUdto", "dVdto", "dSdto", "dUdt", "dVdt", "dSdt")

    def time_step(self, dt, theta, first=False):
        """Start a flow time step (dt <= 0: stationary flow)."""
        _check(lib().rmx_time_step(self.h, dt, theta, int(first)))

    def slip_flow(self):
        """The slip condition at rotated nodes that ends the flow cycle (GRID::SetSlipFlow)."""
        _check(lib().rmx_slip_flow(self.h))

    def time_commit(self):
        _check(lib().rmx_time_commit(self.h))

    def dry_rewet(self):
        """One call of the drying and rewetting (method 2, MODEL::DoDryRewet): (elements
        that got dry, that got wet).  A change of the wet elements renumbers the system."""
        d, w = C.c_int(0), C.c_int(0)
        _check(lib().rmx_dry_rewet(self.h, C.byref(d), C.byref(w)))
        if d.value + w.value:
            self.prepare()
        return d.value, w.value

    def wetdry(self):
        """Per node the reference's dry (64) and marsh (256) bits and those of the previous
        time step (128, 512), per element 1 dry, 2 marsh."""
        nf = np.zeros(self.np, np.int32)
        ef = np.zeros(self.ne, np.int32)
        _check(lib().rmx_get_wetdry(self.h, _p(nf, _i), _p(ef, _i)))
        return nf, ef

    def relax_times(self):
        """(dt_UV, dt_H, dt) of the running flow cycle."""
        a, b, t = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(lib().rmx_get_relax_time(self.h, C.byref(a), C.byref(b), C.byref(t)))
        return a.value, b.value, t.value

    def set_inflow(self, Q, S_in, hydrograph=True):
        """QINLET discharge and reference level for the following Newton steps (from a
        hydrograph: the line is QTINLET)."""
        _check(lib().rmx_set_inflow(self.h, Q, S_in, int(hydrograph)))

    def time_state(self):
        out = {k: np.zeros(self.np) for k in self.TIME_KEYS}
        ptrs = (_d * 9)(*[_p(out[k], _d) for k in self.TIME_KEYS])
        _check(lib().rmx_time_state(self.h, 0, ptrs))
        return out

    def set_time_state(self, **kw):
        arrs = [None if kw.get(k) is None else _c(kw[k], np.float64) for k in self.TIME_KEYS]
        ptrs = (_d * 9)(*[None if a is None else _p(a, _d) for a in arrs])
        _check(lib().rmx_time_state(self.h, 1, ptrs))

    def turbulence(self):
        """Friction and the algebraic model's nodal vt, uu of the current state."""
        vt, uu = np.zeros(self.np), np.zeros(self.np)
        _check(lib().rmx_turbulence(self.h, _p(vt, _d), _p(uu, _d)))
        return vt, uu

    def state_upload(self, U, V, S):
        U, V, S = _c(U, np.float64), _c(V, np.float64), _c(S, np.float64)
        _check(lib().rmx_state_upload(self.h, _p(U, _d), _p(V, _d), _p(S, _d)))

    def state(self):
        U, V, S, z = (np.zeros(self.np) for _ in range(4))
        _check(lib().rmx_state_download(self.h, _p(U, _d), _p(V, _d), _p(S, _d), _p(z, _d)))
        return U, V, S, z

    def prepare(self):
        n = C.c_int(0)
        nnz = C.c_longlong(0)
        _check(lib().rmx_prepare(self.h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value
        return self.n

    def numbering(self):
        eq = np.zeros(3 * self.np, np.int32)
        _check(lib().rmx_get_numbering(self.h, _p(eq, _i)))
        return eq.reshape(self.np, 3)

    def pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        cols = np.zeros(self.nnz, np.int32)
        _check(lib().rmx_get_pattern(self.h, _p(rp, _i), _p(cols, _i)))
        return rp, cols

    def nodal(self):
        cf, bcU = np.zeros(self.np), np.zeros(self.np)
        kind = np.zeros(self.np, np.int32)
        rotm = np.zeros((self.np, 4))
        _check(lib().rmx_get_nodal(self.h, _p(cf, _d), _p(bcU, _d), _p(kind, _i), _p(rotm, _d)))
        return dict(cf=cf, bcU=bcU, kind=kind, rotm=rotm)

    def n_model_elems(self):
        n = C.c_int(0)
        _check(lib().rmx_n_model_elems(self.h, C.byref(n)))
        return n.value

    def element_coefs(self, first, count):
        K = np.zeros((count, 24, 24))
        F = np.zeros((count, 24))
        _check(lib().rmx_element_coefs(self.h, first, count, _p(K, _d), _p(F, _d)))
        return K, F

    # ---- drop-in group -----------------------------------------------------------------
    def assemble(self, prec=F64):
        _check(lib().rmx_assemble(self.h, prec))

    def matrix(self, prec=F64):
        A = np.zeros(self.nnz, np.float32 if prec == F32 else np.float64)
        B = np.zeros(self.n)
        _check(lib().rmx_get_matrix(self.h, prec, A.ctypes.data_as(_v), _p(B, _d)))
        return A, B

    def factors(self, prec=F64):
        L = np.zeros(self.nnz, np.float32 if prec == F32 else np.float64)
        _check(lib().rmx_get_factors(self.h, prec, L.ctypes.data_as(_v)))
        return L

    def scale_l2(self):
        s = C.c_double(0)
        _check(lib().rmx_scale_l2(self.h, C.byref(s)))
        return s.value

    def convert(self, to_prec):
        _check(lib().rmx_convert_matrix(self.h, to_prec))

    def ilu0_factor(self, prec=F64):
        _check(lib().rmx_ilu0_factor(self.h, prec))

    def ilu0_apply(self, b, prec=F64):
        b = _c(b, np.float64)
        x = np.zeros(self.n)
        _check(lib().rmx_ilu0_apply(self.h, prec, _p(b, _d), _p(x, _d)))
        return x

    def silu0_factor(self, prec=F64):
        _check(lib().rmx_silu0_factor(self.h, prec))

    def silu0_apply(self, b, prec=F64):
        b = _c(b, np.float64)
        x = np.zeros(self.n)
        _check(lib().rmx_silu0_apply(self.h, prec, _p(b, _d), _p(x, _d)))
        return x

    def set_strips(self, strips, segments=0):
        """Strips (0: two per CU) and segments per strip (0: blocks of about 560 rows)."""
        _check(lib().rmx_set_strips(self.h, int(strips), int(segments)))

    def sweep_info(self, kind=PREC_SILU0):
        si = SweepInfo()
        _check(lib().rmx_get_sweep_info(self.h, kind, C.byref(si)))
        return si

    def sweep_entries(self, kind=PREC_SILU0):
        si = self.sweep_info(kind)
        return si.lower_entries, si.upper_entries

    def set_coarse(self, segments):
        _check(lib().rmx_set_coarse(self.h, int(segments)))

    def coarse_factor(self, prec=F64):
        _check(lib().rmx_coarse_factor(self.h, prec))

    def aggregates(self):
        a = np.zeros(self.n, np.int32)
        nc = C.c_int(0)
        _check(lib().rmx_get_aggregates(self.h, _p(a, _i), C.byref(nc)))
        return a, nc.value

    def precond_apply(self, kind, b, prec=F64):
        b = _c(b, np.float64)
        x = np.zeros(self.n)
        _check(lib().rmx_precond_apply(self.h, kind, prec, _p(b, _d), _p(x, _d)))
        return x

    # ---- learned preconditioner ----------------------------------------------------------
    @staticmethod
    def mpnn_n_params():
        n = C.c_int(0)
        _check(lib().rmx_mpnn_n_params(C.byref(n)))
        return n.value

    def mpnn_set_weights(self, theta):
        theta = _c(theta, np.float32)
        _check(lib().rmx_mpnn_set_weights(self.h, _p(theta, _f), len(theta)))

    def mpnn_get_weights(self):
        theta = np.zeros(self.mpnn_n_params(), np.float32)
        _check(lib().rmx_mpnn_get_weights(self.h, _p(theta, _f), len(theta)))
        return theta

    def mpnn_forward(self, r):
        """Network correction RMS(r) * network(r) as an equation vector, and the node
        features [np, 8]."""
        r = _c(r, np.float64)
        z = np.zeros(self.n)
        feat = np.zeros((self.np, 8), np.float32)
        _check(lib().rmx_mpnn_forward(self.h, _p(r, _d), _p(z, _d), _p(feat, _f)))
        return z, feat

    def mpnn_add_sample(self, r, capacity=0):
        r = _c(r, np.float64)
        k = C.c_int(0)
        _check(lib().rmx_mpnn_add_sample(self.h, _p(r, _d), int(capacity), C.byref(k)))
        return k.value

    def mpnn_loss_grad(self, sample):
        loss = C.c_double(0)
        g = np.zeros(self.mpnn_n_params(), np.float32)
        _check(lib().rmx_mpnn_loss_grad(self.h, int(sample), C.byref(loss), _p(g, _f)))
        return loss.value, g

    def mpnn_train(self, epochs, lr):
        loss = C.c_double(0)
        _check(lib().rmx_mpnn_train(self.h, int(epochs), float(lr), C.byref(loss)))
        return loss.value

    def blocks(self):
        b = np.zeros(self.n, np.int32)
        _check(lib().rmx_get_blocks(self.h, _p(b, _i)))
        return b

    def spmv(self, x, prec=F64):
        x = _c(x, np.float64)
        r = np.zeros(self.n)
        _check(lib().rmx_spmv(self.h, prec, _p(x, _d), _p(r, _d)))
        return r

    def solve(self, method=BICGSTAB, prec=F64, precond=PREC_ILU0, max_iter=1000, mkyrl=30,
              eps=1e-10, refine=False, refine_tol=1e-8, refine_max=10, proceed=0, next=None):
        o = krylov_opts(method, prec, precond, max_iter, mkyrl, eps, refine, refine_tol,
                        refine_max, proceed, next)
        out = KrylovOut()
        _check(lib().rmx_solve(self.h, C.byref(o), C.byref(out)))
        return out

    def solution(self):
        X = np.zeros(self.n)
        _check(lib().rmx_get_solution(self.h, _p(X, _d)))
        return X

    def set_solution(self, X):
        """The correction the next newton_update applies."""
        X = _c(X, np.float64)
        assert len(X) == self.n
        _check(lib().rmx_set_solution(self.h, _p(X, _d)))

    def set_rhs(self, B):
        B = _c(B, np.float64)
        _check(lib().rmx_set_rhs(self.h, _p(B, _d)))

    def csr_upload(self, rowptr, cols, vals, B):
        rowptr, cols = _c(rowptr, np.int32), _c(cols, np.int32)
        vals, B = _c(vals, np.float64), _c(B, np.float64)
        _check(lib().rmx_csr_upload(self.h, len(B), _p(rowptr, _i), _p(cols, _i), _p(vals, _d),
                                    _p(B, _d)))
        self.n, self.nnz = len(B), int(rowptr[-1])

    def newton_update(self):
        st = StepStats()
        _check(lib().rmx_newton_update(self.h, C.byref(st)))
        return st

    def newton_step(self, method=BICGSTAB, prec=F64, precond=PREC_ILU0, max_iter=1000, mkyrl=30,
                    eps=1e-10, refine=False, refine_tol=1e-8, refine_max=10, proceed=0, next=None):
        o = krylov_opts(method, prec, precond, max_iter, mkyrl, eps, refine, refine_tol,
                        refine_max, proceed, next)
        st = StepStats()
        _check(lib().rmx_newton_step(self.h, C.byref(o), C.byref(st)))
        if not self.n:
            self.prepare_sizes()
        return st

    def prepare_sizes(self):
        return self.prepare()

    def kernel_stats(self, on=None):
        """Launch timing per kernel class: on=True resets and starts it, on=False stops it;
        returns {class: (seconds, algorithmic bytes, launches)} summed since the reset."""
        if on is not None:
            _check(lib().rmx_kernel_stats_enable(self.h, int(bool(on))))
        ks = KernelStats()
        _check(lib().rmx_kernel_stats_get(self.h, C.byref(ks)))
        return {name: (ks.seconds[k], ks.bytes[k], ks.launches[k])
                for k, name in enumerate(KERNEL_CLASSES)}

    def profile_kernels(self, prec=F64, reps=10):
        kt = KernelTimes()
        _check(lib().rmx_profile_kernels(self.h, prec, reps, C.byref(kt)))
        return kt


def hydrograph(tQS, time):
    """Q, S of a hydrograph (rows t, Q, S) at `time`, as the reference's QTINLET lookup."""
    a = np.ascontiguousarray(np.asarray(tQS, np.float64).T)
    q, s_ = C.c_double(), C.c_double()
    _check(lib().rmx_hydrograph(a.shape[1], _p(a[0], _d), _p(a[1], _d), _p(a[2], _d), time,
                                C.byref(q), C.byref(s_)))
    return q.value, s_.value
