"""Python host mirror of the vertex-centred (median-dual) hot path over the
C-ABI (include/ufx.h).

:class:`NCMesh` wraps the preprocessing (the reference's ``geometric``,
NCFVM/geometric.f90:1-46).  The library is the same ``libufx.so`` as the
cell-centred path; there is no fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import nc_mesh
from .cc import UfxError, _check, _d, _i, _sig, _vp, lib  # noqa: F401

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

_sig("ufx_nc_geometric", [C.c_int] * 3 + [_dp, _dp, _ip, _dp, _dp, _dp] + [_ip] * 5
     + [C.POINTER(_vp)])

NC_MESH_I32 = ["fc_p1", "fc_p2", "fc_in", "fc_out", "fc_bc", "pt_bc", "pt_nv2v", "pt_v2v",
               "pt_nv2f", "pt_v2f", "oldnum", "counts"]
NC_MESH_F64 = ["fc_nx", "fc_ny", "fc_area", "fc_cenx", "fc_ceny", "pt_x", "pt_y", "pt_cv",
               "pt_wx", "pt_wy", "cell_cenx", "cell_ceny", "cell_cv"]


class NCMesh:
    """A preprocessed vertex-centred mesh (reference state after ``geometric``)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def from_raw(cls, raw: nc_mesh.NCRawMesh) -> "NCMesh":
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        k = dict(px=f64(raw.x), py=f64(raw.y), pbc=i32(raw.pbc), cx=f64(raw.cen[:, 0]),
                 cy=f64(raw.cen[:, 1]), cv=f64(raw.area), a=i32(raw.edge[:, 0]),
                 b=i32(raw.edge[:, 1]), i=i32(raw.edge[:, 2]), o=i32(raw.edge[:, 3]),
                 bc=i32(raw.ebc))
        h = _vp()
        _check(lib.ufx_nc_geometric(raw.nop, raw.noc, raw.nof, _d(k["px"]), _d(k["py"]),
                                    _i(k["pbc"]), _d(k["cx"]), _d(k["cy"]), _d(k["cv"]),
                                    _i(k["a"]), _i(k["b"]), _i(k["i"]), _i(k["o"]), _i(k["bc"]),
                                    C.byref(h)))
        return cls(h)

    def get(self, name: str) -> np.ndarray:
        n = C.c_int64()
        _check(lib.ufx_mesh_size(self._h, name.encode(), C.byref(n)))
        dt = np.int32 if name in NC_MESH_I32 else np.float64
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
