"""Python host mirror of the vertex-centred (median-dual) hot path over the
C-ABI (include/ufx.h).

:class:`NCMesh` wraps the preprocessing (the reference's ``geometric``,
NCFVM/geometric.f90:1-46); :class:`NCSolver` owns a device context with one
method per reference routine (``save_old``, ``fvresidual``, ``rk_stage``,
``lusgs``, ``residue``) and ``iterate`` for whole iterations of
NCFVM/main.f90:51-88.  The library is the same ``libufx.so`` as the
cell-centred path; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import nc_mesh
from .cc import UfxError, _check, _d, _i, _sig, _vp, lib  # noqa: F401

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

_sig("ufx_nc_geometric", [C.c_int] * 3 + [_dp, _dp, _ip, _dp, _dp, _dp] + [_ip] * 5
     + [C.POINTER(_vp)])
_sig("ufx_nc_geometric_device", [C.c_int] * 3 + [_dp, _dp, _ip, _dp, _dp, _dp] + [_ip] * 5
     + [C.POINTER(_vp)])


class NCParams(C.Structure):
    _fields_ = [("mach", C.c_double), ("aoa_deg", C.c_double), ("cfl_max", C.c_double),
                ("grad", C.c_int), ("timemode", C.c_int), ("precision", C.c_int),
                ("flux", C.c_int), ("limiter", C.c_int), ("lusgs_order", C.c_int)]


_sig("ufx_nc_create", [_vp, C.POINTER(NCParams), C.POINTER(_vp)])
_sig("ufx_nc_destroy", [_vp])
_sig("ufx_nc_save_old", [_vp])
_sig("ufx_nc_fvresidual", [_vp, C.c_double])
_sig("ufx_nc_rk_stage", [_vp, C.c_int, C.c_double])
_sig("ufx_nc_lusgs", [_vp, C.c_double])
_sig("ufx_nc_residue", [_vp, C.c_int, _dp])
_sig("ufx_nc_iterate", [_vp, C.c_int, C.c_int, _dp])
_sig("ufx_nc_set", [_vp, C.c_char_p, _dp])
_sig("ufx_nc_get", [_vp, C.c_char_p, _dp])
_sig("ufx_nc_sync", [_vp])
_sig("ufx_nc_profile", [_vp, C.c_int])
_sig("ufx_nc_profile_read", [_vp, _dp, _ip])
_sig("ufx_nc_forces", [_vp, _dp, _dp])

NC_EXPORTED = ["ufx_nc_geometric", "ufx_nc_geometric_device", "ufx_nc_create", "ufx_nc_destroy", "ufx_nc_save_old",
               "ufx_nc_fvresidual", "ufx_nc_rk_stage", "ufx_nc_lusgs", "ufx_nc_residue",
               "ufx_nc_iterate", "ufx_nc_set", "ufx_nc_get", "ufx_nc_sync", "ufx_nc_profile",
               "ufx_nc_profile_read", "ufx_nc_forces"]
NKCAT = ["save_old", "gradient", "residual", "norm", "lusgs"]

NC_MESH_I32 = ["fc_p1", "fc_p2", "fc_in", "fc_out", "fc_bc", "pt_bc", "pt_nv2v", "pt_v2v",
               "pt_nv2f", "pt_v2f", "oldnum", "counts"]
NC_MESH_F64 = ["fc_nx", "fc_ny", "fc_area", "fc_cenx", "fc_ceny", "pt_x", "pt_y", "pt_cv",
               "pt_wx", "pt_wy", "cell_cenx", "cell_ceny", "cell_cv"]


class NCMesh:
    """A preprocessed vertex-centred mesh (reference state after ``geometric``)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def from_raw(cls, raw: nc_mesh.NCRawMesh, device: bool = False) -> "NCMesh":
        """Preprocess (NCFVM/geometric.f90:1-46); with ``device`` the per-edge
        and per-vertex passes run on the GPU (the same arrays bit for bit)."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        k = dict(px=f64(raw.x), py=f64(raw.y), pbc=i32(raw.pbc), cx=f64(raw.cen[:, 0]),
                 cy=f64(raw.cen[:, 1]), cv=f64(raw.area), a=i32(raw.edge[:, 0]),
                 b=i32(raw.edge[:, 1]), i=i32(raw.edge[:, 2]), o=i32(raw.edge[:, 3]),
                 bc=i32(raw.ebc))
        h = _vp()
        fn = lib.ufx_nc_geometric_device if device else lib.ufx_nc_geometric
        _check(fn(raw.nop, raw.noc, raw.nof, _d(k["px"]), _d(k["py"]), _i(k["pbc"]), _d(k["cx"]),
                  _d(k["cy"]), _d(k["cv"]), _i(k["a"]), _i(k["b"]), _i(k["i"]), _i(k["o"]),
                  _i(k["bc"]), C.byref(h)))
        return cls(h)

    def get(self, name: str) -> np.ndarray:
        n = C.c_int64()
        _check(lib.ufx_mesh_size(self._h, name.encode(), C.byref(n)))
        dt = np.int32 if name in NC_MESH_I32 or name == "pt_colour" else np.float64
        out = np.empty(n.value, dtype=dt)
        _check(lib.ufx_mesh_get(self._h, name.encode(), out.ctypes.data_as(_vp)))
        return out

    def arrays(self) -> dict:
        """All arrays, indices 1-based (the reference's numbering)."""
        return {n: self.get(n) for n in NC_MESH_I32 + NC_MESH_F64}

    @property
    def counts(self):
        nop, noc, nof = (int(v) for v in self.get("counts"))
        return dict(nop=nop, noc=noc, nof=nof)

    def close(self):
        if self._h:
            lib.ufx_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NCSolver:
    """A device context of the vertex-centred solver (FP64)."""

    SHAPES = {"qc": 4, "qold": 4, "qp": 4, "res": 4, "phi": 4, "grad": 8, "la": 1, "dt": 1}

    def __init__(self, mesh: NCMesh, *, mach=0.4, aoa_deg=0.0, cfl_max=50.0, grad="gg",
                 timemode="lusgs", flux="vanleer", limiter="vkn", lusgs_order="reference",
                 precision="fp64"):
        self.mesh = mesh
        self.counts = mesh.counts
        self.timemode = timemode
        self.cfl_max = cfl_max
        prm = NCParams(mach, aoa_deg, cfl_max, {"lsqr": 1, "gg": 2}[grad],
                       {"rk3": 1, "lusgs": 2}[timemode], {"fp64": 0, "mixed": 1}[precision],
                       {"ausm": 2, "vanleer": 3}[flux],
                       {"vkn": 0, "bj": 1}[limiter],
                       {"reference": 0, "multicolour": 1}[lusgs_order])
        self._h = _vp()
        _check(lib.ufx_nc_create(mesh._h, C.byref(prm), C.byref(self._h)))
        self.iter = 0

    def cfl(self, it):
        return 0.5 * self.cfl_max * (1.0 + math.tanh(it * 7.0 / 500.0 - 3.5))

    def save_old(self): _check(lib.ufx_nc_save_old(self._h))
    def fvresidual(self, cfl): _check(lib.ufx_nc_fvresidual(self._h, cfl))
    def rk_stage(self, k, cfl): _check(lib.ufx_nc_rk_stage(self._h, k, cfl))
    def lusgs(self, cfl): _check(lib.ufx_nc_lusgs(self._h, cfl))
    def sync(self): _check(lib.ufx_nc_sync(self._h))

    def residue(self, it):
        out = np.zeros(5)
        _check(lib.ufx_nc_residue(self._h, it, _d(out)))
        return out

    def iterate(self, n, hist=True):
        h = np.zeros((n, 5)) if hist else None
        _check(lib.ufx_nc_iterate(self._h, self.iter, n, _d(h) if hist else None))
        self.iter += n
        return h

    def forces(self):
        """Wall force of the current state (the wall closure of solid_flux
        summed in edge order; the reference's vertex-centred solver has no
        force routine): (cl, cd, fx, fy) and -Cp at the first end point of
        each wall edge, in edge order."""
        nw = int(np.sum(self.mesh.get("fc_bc") == 1001))
        out = np.zeros(4)
        cp = np.zeros(max(nw, 1))
        _check(lib.ufx_nc_forces(self._h, _d(out), _d(cp)))
        return out, cp[:nw]

    def get(self, name):
        n = self.counts["nof"] if name == "fc_la" else self.counts["nop"]
        w = self.SHAPES.get(name, 1)
        out = np.zeros((n, w) if w > 1 else n)
        _check(lib.ufx_nc_get(self._h, name.encode(), _d(out)))
        return out

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib.ufx_nc_set(self._h, name.encode(), _d(a)))

    def profile(self, enable=True):
        _check(lib.ufx_nc_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        ms = np.zeros(len(NKCAT))
        cnt = np.zeros(len(NKCAT), dtype=np.int32)
        _check(lib.ufx_nc_profile_read(self._h, _d(ms), cnt.ctypes.data_as(_ip)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(NKCAT)}

    def close(self):
        if getattr(self, "_h", None):
            lib.ufx_nc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
