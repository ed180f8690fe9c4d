"""Python host mirror of the cell-centred hot path over the C-ABI (include/ufx.h).

:class:`Mesh` wraps the preprocessing (the reference's ``geometric``);
:class:`Solver` owns a device context and exposes one method per reference
routine (``prim``, ``avg_c2v``, ``time_step``, ``save_old``, ``gradient``,
``limit``, ``fvresidual``, ``rk_update``, ``lusgs``, ``residue``,
``check_positivity``) plus ``iterate`` for whole iterations of the reference's
main loop (CCFVM/main.F90:57-128).  Errors the reference reports with STOP
raise :class:`UfxError` carrying the reference's message.

The library is loaded from the package directory; there is no fallback: if
``libufx.so`` is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import os
import time

import numpy as np

from . import meshgen

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libufx.so")

FLUX = {"roe": 1, "ausm": 2, "vanleer": 3, "rusanov": 4}
GRAD = {"lsqr": 1, "gg": 2, "ggfc": 3}
FLOW = {"inviscid": 1, "laminar": 2}
TIME = {"rk3": 1, "lusgs": 2}
PREC = {"fp64": 0, "mixed": 1}

ERRORS = {1: "argument", 2: "mesh", 3: "hip", 4: "state", 5: "unsupported"}


class UfxError(RuntimeError):
    """A failing C-ABI call; ``code`` is the UFX_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m unsflow2d_amd.build`")
    return C.CDLL(LIB_PATH)


lib = _load()

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class Params(C.Structure):
    _fields_ = [("mach", C.c_double), ("aoa_deg", C.c_double), ("reynolds", C.c_double),
                ("t_wall", C.c_double), ("flow_type", C.c_int), ("iwall", C.c_int),
                ("flux", C.c_int), ("grad", C.c_int), ("ilimit", C.c_int), ("irs", C.c_int),
                ("timemode", C.c_int), ("precision", C.c_int), ("cfl_max", C.c_double),
                ("xref", C.c_double), ("yref", C.c_double), ("lusgs_order", C.c_int)]


def _sig(name, args, res=C.c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("ufx_last_error", [], C.c_char_p)
_sig("ufx_cc_geometric", [C.c_int] * 4 + [_dp] * 5 + [_ip] * 5 + [_ip] * 3 + [_dp] * 2 + [_ip, C.POINTER(_vp)])
_sig("ufx_cc_geometric_device", [C.c_int] * 4 + [_dp] * 5 + [_ip] * 5 + [_ip] * 3 + [_dp] * 2 + [_ip, C.POINTER(_vp)])
_sig("ufx_cc_mesh_import", [C.c_int] * 4 + [_dp] * 2 + [_ip] * 5 + [_dp] * 6 + [_dp] * 7 + [_ip] * 3 + [_ip] * 2 + [_dp, C.POINTER(_vp)])
_sig("ufx_mesh_size", [_vp, C.c_char_p, C.POINTER(C.c_int64)])
_sig("ufx_mesh_get", [_vp, C.c_char_p, _vp])
_sig("ufx_mesh_destroy", [_vp], None)
_sig("ufx_cc_create", [_vp, C.POINTER(Params), C.POINTER(_vp)])
_sig("ufx_cc_destroy", [_vp])
_sig("ufx_cc_set", [_vp, C.c_char_p, _dp])
_sig("ufx_cc_get", [_vp, C.c_char_p, _dp])
for _n in ("prim", "avg_c2v", "save_old", "gradient", "limit", "fvresidual", "lusgs", "sync"):
    _sig(f"ufx_cc_{_n}", [_vp])
_sig("ufx_cc_time_step", [_vp, C.c_double])
_sig("ufx_cc_rk_update", [_vp, C.c_int])
_sig("ufx_cc_rk_stage", [_vp, C.c_int])
_sig("ufx_cc_residue", [_vp, C.c_int, _dp, _dp])
_sig("ufx_cc_check_positivity", [_vp, _ip])
_sig("ufx_cc_error", [_vp, _ip, _ip, _ip])
_sig("ufx_cc_forces", [_vp, _dp, _dp, _dp])
_sig("ufx_cc_iterate", [_vp, C.c_int, C.c_int, _dp, _dp])
_sig("ufx_cc_iterate_graph", [_vp, C.c_int, C.c_int, C.c_double, _dp])
_sig("ufx_cc_keep_residual", [_vp, C.c_int])
_sig("ufx_cc_history", [_vp, C.c_int, _dp])
_sig("ufx_cc_set_stage_form", [_vp, C.c_int, C.c_int])
_sig("ufx_cc_get_stage_form", [_vp, _ip, _ip])
_sig("ufx_cc_residual_norm", [_vp, C.c_int, _dp])
_sig("ufx_cc_residual_norm_noalloc", [_vp, C.c_int, _dp])
_sig("ufx_cc_stream", [_vp], C.c_void_p)
_sig("ufx_cc_profile", [_vp, C.c_int])
_sig("ufx_cc_profile_mask", [_vp, C.c_uint])
_sig("ufx_cc_profile_read", [_vp, _dp, _ip])
_sig("ufx_cc_batch_create", [C.POINTER(_vp), C.c_int, C.POINTER(_vp)])
_sig("ufx_cc_batch_set_active", [_vp, _ip])
_sig("ufx_cc_batch_iterate", [_vp, C.c_int, C.c_double])
_sig("ufx_cc_batch_residual_norms", [_vp, C.c_int, _dp])
_sig("ufx_cc_batch_destroy", [_vp])

KCAT = ["prim", "range", "avg_c2v", "time_step", "gradient", "stage", "norm", "other", "bface"]

EXPORTED = [
    "ufx_last_error", "ufx_cc_geometric", "ufx_cc_geometric_device", "ufx_cc_mesh_import", "ufx_mesh_size", "ufx_mesh_get",
    "ufx_mesh_destroy", "ufx_cc_create", "ufx_cc_destroy", "ufx_cc_set", "ufx_cc_get",
    "ufx_cc_prim", "ufx_cc_avg_c2v", "ufx_cc_time_step", "ufx_cc_save_old", "ufx_cc_gradient",
    "ufx_cc_limit", "ufx_cc_fvresidual", "ufx_cc_rk_update", "ufx_cc_rk_stage", "ufx_cc_lusgs",
    "ufx_cc_residue", "ufx_cc_check_positivity", "ufx_cc_forces", "ufx_cc_iterate",
    "ufx_cc_iterate_graph", "ufx_cc_history", "ufx_cc_keep_residual", "ufx_cc_residual_norm",
    "ufx_cc_residual_norm_noalloc", "ufx_cc_set_stage_form", "ufx_cc_get_stage_form",
    "ufx_cc_profile_mask",
    "ufx_cc_sync", "ufx_cc_error", "ufx_cc_stream", "ufx_cc_profile", "ufx_cc_profile_read",
    "ufx_cc_batch_create", "ufx_cc_batch_set_active", "ufx_cc_batch_iterate",
    "ufx_cc_batch_residual_norms", "ufx_cc_batch_destroy",
]


def _check(rc: int) -> None:
    if rc != 0:
        raise UfxError(rc, lib.ufx_last_error().decode())


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


MESH_I32 = ["fc_p1", "fc_p2", "fc_in", "fc_out", "fc_bc", "cell_nc2c", "cell_nc2f", "cell_c2f",
            "pt_nv2c", "pt_v2c", "oldnum", "counts"]
# the vertex walk round each cell (output only; empty for an imported mesh)
MESH_I32_WALK = ["cell_nc2v", "cell_c2v"]
MESH_F64 = ["fc_nx", "fc_ny", "fc_area", "fc_cenx", "fc_ceny", "fc_cov", "cell_cenx",
            "cell_ceny", "cell_cv", "cell_cov", "cell_r11", "cell_r12", "cell_r22", "pt_x",
            "pt_y", "pt_wt"]


class Mesh:
    """A preprocessed cell-centred mesh (reference state after ``geometric``)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def from_raw(cls, raw: meshgen.RawMesh, device: bool = False) -> "Mesh":
        """Preprocess a geometry.inp-form mesh (CCFVM/geometric.F90:1-57);
        with ``device`` its per-face, per-cell and per-vertex passes run on
        the GPU (the same arrays bit for bit)."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        keep = dict(px=f64(raw.x), py=f64(raw.y), cx=f64(raw.cen[:, 0]), cy=f64(raw.cen[:, 1]),
                    cv=f64(raw.area), a=i32(raw.fint[:, 0]), b=i32(raw.fint[:, 1]),
                    i=i32(raw.fint[:, 2]), o=i32(raw.fint[:, 3]),
                    ba=i32(raw.fbnd[:, 0]), bb=i32(raw.fbnd[:, 1]), bi=i32(raw.fbnd[:, 2]),
                    bnx=f64(raw.fbn[:, 0]), bny=f64(raw.fbn[:, 1]), bc=i32(raw.fbc))
        h = _vp()
        k = keep
        fn = lib.ufx_cc_geometric_device if device else lib.ufx_cc_geometric
        _check(fn(raw.nop, raw.noc, raw.nfi, raw.nbf, _d(k["px"]), _d(k["py"]), _d(k["cx"]),
                  _d(k["cy"]), _d(k["cv"]), _i(k["a"]), _i(k["b"]), _i(k["i"]), _i(k["o"]), None,
                  _i(k["ba"]), _i(k["bb"]), _i(k["bi"]), _d(k["bnx"]), _d(k["bny"]), _i(k["bc"]),
                  C.byref(h)))
        return cls(h)

    @classmethod
    def import_arrays(cls, a: dict, nfi: int) -> "Mesh":
        """Hand over post-``geometric`` arrays as the reference holds them
        (1-based indices, the keys of :meth:`arrays`)."""
        f64 = lambda n: np.ascontiguousarray(a[n], dtype=np.float64)
        i32 = lambda n: np.ascontiguousarray(a[n], dtype=np.int32)
        nop, noc = len(a["pt_x"]), len(a["cell_cv"])
        nof = len(a["fc_p1"])
        nbf = nof - nfi
        nc2f = i32("cell_nc2f")
        c2f_ptr = np.zeros(noc + 1, dtype=np.int32)
        np.cumsum(nc2f, out=c2f_ptr[1:])
        nv2c = i32("pt_nv2c")
        v2c_ptr = np.zeros(nop + 1, dtype=np.int32)
        np.cumsum(nv2c, out=v2c_ptr[1:])
        keep = {n: (i32(n) if n in MESH_I32 else f64(n)) for n in a if n not in ("counts", "oldnum")}
        keep["c2f_ptr"], keep["v2c_ptr"] = c2f_ptr, v2c_ptr
        k = keep
        h = _vp()
        _check(lib.ufx_cc_mesh_import(
            nop, noc, nfi, nbf, _d(k["pt_x"]), _d(k["pt_y"]), _i(k["fc_p1"]), _i(k["fc_p2"]),
            _i(k["fc_in"]), _i(k["fc_out"]), _i(k["fc_bc"]), _d(k["fc_nx"]), _d(k["fc_ny"]),
            _d(k["fc_area"]), _d(k["fc_cenx"]), _d(k["fc_ceny"]), _d(k["fc_cov"]),
            _d(k["cell_cenx"]), _d(k["cell_ceny"]), _d(k["cell_cv"]), _d(k["cell_cov"]),
            _d(k["cell_r11"]), _d(k["cell_r12"]), _d(k["cell_r22"]), _i(k["cell_nc2c"]),
            _i(k["c2f_ptr"]), _i(k["cell_c2f"]), _i(k["v2c_ptr"]), _i(k["pt_v2c"]), _d(k["pt_wt"]),
            C.byref(h)))
        return cls(h)

    def get(self, name: str) -> np.ndarray:
        n = C.c_int64()
        _check(lib.ufx_mesh_size(self._h, name.encode(), C.byref(n)))
        dt = (np.int32 if name in MESH_I32 or name in MESH_I32_WALK or name == "cell_colour"
              else np.float64)
        out = np.empty(n.value, dtype=dt)
        _check(lib.ufx_mesh_get(self._h, name.encode(), out.ctypes.data_as(_vp)))
        return out

    def arrays(self) -> dict:
        """All arrays, indices 1-based (the reference's numbering)."""
        return {n: self.get(n) for n in MESH_I32 + MESH_F64}

    @property
    def counts(self):
        nop, noc, nof, nbf = (int(v) for v in self.get("counts"))
        return dict(nop=nop, noc=noc, nof=nof, nbf=nbf, nfi=nof - nbf, ncell=noc + nbf)

    def close(self):
        if self._h:
            lib.ufx_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclasses.dataclass
class Config:
    """Deck settings (CCFVM/read_input.F90:31-50) the hot path depends on."""
    mach: float = 0.5
    aoa_deg: float = 0.0
    reynolds: float = 1000.0
    t_wall: float = 300.0
    flow_type: str = "inviscid"
    iwall: int = 1
    flux: str = "roe"
    grad: str = "lsqr"
    ilimit: int = 1
    irs: int = 0
    timemode: str = "rk3"
    precision: str = "fp64"
    cfl_type: str = "ramp"
    cfl_min: float = 0.8
    cfl_max: float = 0.8
    cfl_ramp_steps: int = 1
    xref: float = 0.0
    yref: float = 0.0
    lusgs_order: str = "reference"

    def params(self) -> Params:
        return Params(self.mach, self.aoa_deg, self.reynolds, self.t_wall, FLOW[self.flow_type],
                      self.iwall, FLUX[self.flux], GRAD[self.grad], self.ilimit, self.irs,
                      TIME[self.timemode], PREC[self.precision], self.cfl_max, self.xref, self.yref,
                      {"reference": 0, "multicolour": 1}[self.lusgs_order])

    def cfl(self, it: int, previous: float = float("nan")) -> float:
        """CFL of iteration ``it``, CCFVM/main.F90:148-168 (the ramp holds its
        last value after ``cfl_ramp_steps``)."""
        if self.cfl_type == "tanh":
            s = float(it) / float(self.cfl_ramp_steps - 1)
            return 0.5 * self.cfl_max * (1.0 + math.tanh(s * 7.0 - 3.5))
        if it <= self.cfl_ramp_steps:
            s = float(it) / float(self.cfl_ramp_steps)
            return self.cfl_min + (self.cfl_max - self.cfl_min) * (1.0 - math.exp(-s * 0.5)) / (
                1.0 - math.exp(-0.5))
        return previous


class Solver:
    """A device context; methods mirror the reference routines."""

    def __init__(self, mesh: Mesh, cfg: Config):
        self.mesh = mesh
        self.cfg = cfg
        self.counts = mesh.counts
        self._h = _vp()
        self._p = cfg.params()
        _check(lib.ufx_cc_create(mesh._h, C.byref(self._p), C.byref(self._h)))
        self.iter = 0
        self.cfl = float("nan")

    NVAR = {"qc": 4, "qp": 5, "qold": 4, "res": 4, "phi": 4, "grad": 10, "dt": 1, "la": 1,
            "fc_la": 1, "fc_mu": 1, "cell_mul": 1, "pt_prim": 5, "pt_mu": 1,
            "fc_grad": 10}

    def _count(self, name):
        c = self.counts
        if name in ("qc", "qp", "cell_mul"):
            return c["ncell"]
        if name.startswith("fc_"):
            return c["nof"]
        if name.startswith("pt_"):
            return c["nop"]
        return c["noc"]

    def get(self, name: str) -> np.ndarray:
        n, k = self._count(name), self.NVAR[name]
        out = np.empty((n, k), dtype=np.float64)
        _check(lib.ufx_cc_get(self._h, name.encode(), _d(out)))
        return out if k > 1 else out[:, 0]

    def set(self, name: str, arr) -> None:
        n, k = self._count(name), self.NVAR[name]
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(n, k))
        _check(lib.ufx_cc_set(self._h, name.encode(), _d(a)))

    # the reference routines
    def prim(self): _check(lib.ufx_cc_prim(self._h))
    def avg_c2v(self): _check(lib.ufx_cc_avg_c2v(self._h))
    def time_step(self, cfl: float): _check(lib.ufx_cc_time_step(self._h, cfl))
    def save_old(self): _check(lib.ufx_cc_save_old(self._h))
    def gradient(self): _check(lib.ufx_cc_gradient(self._h))
    def limit(self): _check(lib.ufx_cc_limit(self._h))
    def fvresidual(self): _check(lib.ufx_cc_fvresidual(self._h))
    def rk_update(self, stage: int): _check(lib.ufx_cc_rk_update(self._h, stage))
    def rk_stage(self, stage: int): _check(lib.ufx_cc_rk_stage(self._h, stage))
    def lusgs(self): _check(lib.ufx_cc_lusgs(self._h))
    def sync(self): _check(lib.ufx_cc_sync(self._h))

    def residue(self, it: int):
        f, raw = C.c_double(), C.c_double()
        _check(lib.ufx_cc_residue(self._h, it, C.byref(f), C.byref(raw)))
        return f.value, raw.value

    ERROR_KIND = {0: None, 1: "roe", 2: "positivity", 3: "area"}

    def error(self) -> dict:
        """The context's device-detected error (``ufx_cc_error``): kind
        (None, "roe", "positivity", "area") and the 1-based reference face
        and cell (0 where they do not apply)."""
        k, f, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib.ufx_cc_error(self._h, C.byref(k), C.byref(f), C.byref(c)))
        return dict(kind=self.ERROR_KIND.get(k.value, k.value), face=f.value, cell=c.value)

    def check_positivity(self) -> int:
        bad = C.c_int()
        rc = lib.ufx_cc_check_positivity(self._h, C.byref(bad))
        if rc not in (0, 4):
            _check(rc)
        return bad.value

    def forces(self, wall: bool = False):
        """find_cl_cd (CCFVM/main.F90:263-355).  With ``wall`` also -Cp and Cf
        per wall face (bc 1001) in boundary order, under "mcp" and "cf"."""
        out = np.zeros(8)
        mcp = cf = None
        if wall:
            nw = int(np.sum(self.mesh.get("fc_bc") == 1001))
            mcp, cf = np.zeros(max(nw, 1)), np.zeros(max(nw, 1))
        _check(lib.ufx_cc_forces(self._h, _d(out), _d(mcp) if wall else None,
                                 _d(cf) if wall else None))
        keys = ["cl", "cd", "cl1", "cd1", "cm", "fx1", "fy1", "dtglobal"]
        res = dict(zip(keys, out.tolist()))
        if wall:
            res["mcp"], res["cf"] = mcp[:nw], cf[:nw]
        return res

    def iterate(self, niter: int, hist: bool = True):
        """Run ``niter`` iterations of the reference's main loop on the device."""
        cfls = np.empty(niter)
        c = self.cfl
        for k in range(niter):
            c = self.cfg.cfl(self.iter + 1 + k, c)
            cfls[k] = c
        h = np.empty(niter) if hist else None
        _check(lib.ufx_cc_iterate(self._h, self.iter, niter, _d(cfls), _d(h) if hist else None))
        self.iter += niter
        self.cfl = c
        return h

    def iterate_graph(self, niter: int, hist: bool = True):
        """``niter`` iterations replayed from a captured HIP graph (from
        iteration 2 on, while the CFL stays constant); the same results as
        :meth:`iterate`."""
        if self.iter < 1:
            raise ValueError("iterate_graph starts at iteration 2: run iterate(1) first")
        c = self.cfl
        for k in range(niter):
            c2 = self.cfg.cfl(self.iter + 1 + k, c)
            if k and c2 != c:
                raise ValueError("iterate_graph needs a constant CFL over the chunk")
            c = c2
        h = np.empty(niter) if hist else None
        _check(lib.ufx_cc_iterate_graph(self._h, self.iter, niter, c, _d(h) if hist else None))
        self.iter += niter
        self.cfl = c
        return h

    def history(self, n: int) -> np.ndarray:
        """The normalised residuals of the last whole-iteration call (also a
        batched one), one per iteration."""
        h = np.empty(n)
        _check(lib.ufx_cc_history(self._h, n, _d(h)))
        return h

    STAGE_FORMS = {"gather": (0, 0), "gather_lds": (0, 1), "shared": (1, 0),
                   "shared_lds": (1, 1)}

    def set_stage_form(self, form: str) -> None:
        """Residual-stage kernel form: "gather" (each cell its own faces),
        "shared" (each interior flux once per block, an edge batch in LDS),
        with "_lds" the block's records staged in LDS."""
        sh, lds = self.STAGE_FORMS[form]
        _check(lib.ufx_cc_set_stage_form(self._h, sh, lds))

    @property
    def stage_form(self) -> str:
        sh, lds = C.c_int(), C.c_int()
        _check(lib.ufx_cc_get_stage_form(self._h, C.byref(sh), C.byref(lds)))
        return {v: k for k, v in self.STAGE_FORMS.items()}[(sh.value, lds.value)]

    def tune_stage_form(self, iters: int = 8, reps: int = 3, log=None) -> dict:
        """Pick the residual-stage form that runs whole iterations of this
        context fastest on this GPU: each form runs one untimed and
        ``reps - 1`` timed batches of ``iters`` iterations (host clock,
        synchronised).  The tuning iterations are real iterations: they
        advance the state and the iteration count.  The fastest form is left
        set; returns it with every form's milliseconds per iteration (None
        where the form does not fit in LDS on this mesh)."""
        times = {}
        for form in self.STAGE_FORMS:
            try:
                self.set_stage_form(form)
            except UfxError:
                times[form] = None
                continue
            self.iterate(iters, hist=False)
            self.sync()
            t0 = time.perf_counter()
            for _ in range(reps - 1):
                self.iterate(iters, hist=False)
            self.sync()
            times[form] = (time.perf_counter() - t0) * 1e3 / (iters * max(1, reps - 1))
            if log:
                log(f"stage form {form}: {times[form]:.4f} ms/iteration")
        best = min((f for f in times if times[f] is not None), key=lambda f: times[f])
        self.set_stage_form(best)
        return {"form": best, "ms_per_iteration": times}

    def keep_residual(self, on: bool = True) -> None:
        """Whole iterations also store the last stage's residual."""
        _check(lib.ufx_cc_keep_residual(self._h, 1 if on else 0))

    def residual_norm(self, var: int = -1) -> float:
        """L2 norm of the stored residual: one variable (0 E, 1 rho, 2 rho u,
        3 rho v) or all four (-1)."""
        out = C.c_double()
        _check(lib.ufx_cc_residual_norm(self._h, var, C.byref(out)))
        return out.value

    def profile(self, enable: bool = True) -> None:
        """Start (or stop) per-kernel HIP-event timing on the context's stream."""
        _check(lib.ufx_cc_profile(self._h, 1 if enable else 0))

    def profile_only(self, *categories: str) -> None:
        """Time only the named kernel categories (see KCAT); none stops."""
        mask = 0
        for c in categories:
            mask |= 1 << KCAT.index(c)
        _check(lib.ufx_cc_profile_mask(self._h, mask))

    def profile_read(self) -> dict:
        """{category: (milliseconds, launches)} since the last read."""
        ms = np.zeros(len(KCAT))
        cnt = np.zeros(len(KCAT), dtype=np.int32)
        _check(lib.ufx_cc_profile_read(self._h, _d(ms), _i(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KCAT)}

    @property
    def stream(self) -> int:
        return lib.ufx_cc_stream(self._h)

    def close(self):
        if self._h:
            lib.ufx_cc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many solvers advanced together, one launch per kernel for all of them
    (``ufx_cc_batch_*``): inviscid RK3 cases with equal scheme options, each
    past iteration 1.  Per cell the result equals each solver's own
    :meth:`Solver.iterate_graph`."""

    def __init__(self, solvers: list):
        self.solvers = list(solvers)
        arr = (_vp * len(self.solvers))(*[s._h for s in self.solvers])
        self._h = _vp()
        _check(lib.ufx_cc_batch_create(arr, len(self.solvers), C.byref(self._h)))
        self.active = np.ones(len(self.solvers), dtype=np.int32)

    def iterate(self, niter: int):
        """``niter`` iterations of every active solver at its (constant) CFL."""
        live = [s for s, a in zip(self.solvers, self.active) if a]
        if not live:
            return
        cfls = set()
        for s in live:
            if s.iter < 1:
                raise ValueError("batched iterations start at iteration 2: run iterate(1) first")
            c = s.cfl
            for k in range(niter):
                c = s.cfg.cfl(s.iter + 1 + k, c)
                cfls.add(c)
        if len(cfls) != 1:
            raise ValueError("batched iterations need one constant CFL for all cases")
        c = cfls.pop()
        _check(lib.ufx_cc_batch_iterate(self._h, niter, c))
        for s in live:
            s.cfl = c
            s.iter += niter

    def residual_norms(self, var: int = -1) -> np.ndarray:
        """Each solver's :meth:`Solver.residual_norm` (``var``) from one
        launch; NaN for retired solvers."""
        out = np.empty(len(self.solvers))
        _check(lib.ufx_cc_batch_residual_norms(self._h, var, _d(out)))
        return out

    def retire(self, k: int):
        """Take solver k out of later iterations (it may be closed after)."""
        self.active[k] = 0
        _check(lib.ufx_cc_batch_set_active(self._h, _i(self.active)))

    def close(self):
        if self._h:
            lib.ufx_cc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
