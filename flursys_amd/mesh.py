"""Unstructured 2D meshes and their Poisson systems built on the GPU.

DeviceMesh2D mirrors UnstructuredMesh::from_cells (src/mesh.rs:317-423) for
2D polygon cells, and assemble_poisson (src/poisson.rs:146-263, with
apply_reference :303-337) returns a CsrMatrix that stays on the device for
the CSR solvers of flursys_amd.linear -- the reference's setup path
(CsrBuilder, HashMap face lookup) run as GPU sorts and scans
(include/flursys_hip.h, fl_mesh2d_*).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, load, ptr
from .linear import CsrMatrix

# MeshError / PoissonError variant names (src/mesh.rs:270-310, src/poisson.rs:78-88)
MESH_KINDS = {1: "EmptyMesh", 2: "NonFinitePoint", 3: "InvalidCellDefinition",
              4: "InvalidPointIndex", 5: "DegenerateFace", 6: "NonManifoldFace",
              7: "ZeroVolumeCell", 8: "InconsistentFaceOrientation"}
POISSON_KINDS = {1: "MissingReferenceForPureNeumann", 2: "IncompatiblePureNeumannSource",
                 3: "InvalidReferenceCell", 4: "InvalidReferenceValue", 5: "NonFiniteSource",
                 6: "DegenerateOwnerNeighbourDistance", 7: "InvalidNonOrthogonalGeometry",
                 8: "DegenerateOwnerFaceDistance", 9: "MissingBoundaryCondition",
                 10: "Numerics", 11: "Linear"}
ZERO_GRADIENT, FIXED_VALUE = 0, 1


class MeshError(ValueError):
    def __init__(self, kind, message):
        super().__init__(f"{kind}: {message}")
        self.kind = kind


class PoissonError(ValueError):
    def __init__(self, kind, message, detail=0.0):
        super().__init__(f"{kind}: {message}")
        self.kind, self.detail = kind, detail


def _last_error(ctx):
    raw = load().fl_last_error(ctx.handle)
    return raw.decode() if raw else ""


class DeviceMesh2D:
    """A 2D polygon mesh (3..16 vertices per cell) built and kept on the GPU."""

    def __init__(self, points, cells, ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        xy = np.ascontiguousarray(np.asarray(points, dtype=np.float64)[:, :2])
        if isinstance(cells, np.ndarray) and cells.ndim == 2:
            vtx = np.ascontiguousarray(cells, dtype=np.int64).ravel()
            off = np.arange(0, vtx.size + 1, cells.shape[1], dtype=np.int64)
        else:
            off = np.zeros(len(cells) + 1, np.int64)
            off[1:] = np.cumsum([len(c) for c in cells])
            vtx = np.ascontiguousarray(np.concatenate([np.asarray(c, np.int64) for c in cells])
                                       if len(cells) else np.zeros(1, np.int64))
        h = C.c_void_p()
        kind = C.c_int32()
        st = load().fl_mesh2d_create(self.ctx.handle, len(xy), ptr(xy), len(off) - 1, ptr(off),
                                     ptr(vtx), C.byref(h), C.byref(kind))
        if st == 14:
            raise MeshError(MESH_KINDS.get(kind.value, str(kind.value)), _last_error(self.ctx))
        check(st, self.ctx.handle)
        self.handle = h
        nc, nf = C.c_int64(), C.c_int64()
        check(load().fl_mesh2d_info(h, C.byref(nc), C.byref(nf)), self.ctx.handle)
        self.ncells, self.nfaces = nc.value, nf.value

    def close(self):
        if getattr(self, "handle", None):
            load().fl_mesh2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        """Cell centres and areas, face owner / neighbour, centres, area vectors."""
        nc, nf = self.ncells, self.nfaces
        g = {"cx": np.empty(nc), "cy": np.empty(nc), "volume": np.empty(nc),
             "owner": np.empty(nf, np.int64), "neighbour": np.empty(nf, np.int64),
             "fcx": np.empty(nf), "fcy": np.empty(nf), "sx": np.empty(nf), "sy": np.empty(nf)}
        check(load().fl_mesh2d_get(self.handle, *[ptr(g[k]) for k in (
            "cx", "cy", "volume", "owner", "neighbour", "fcx", "fcy", "sx", "sy")]),
            self.ctx.handle)
        return g

    def assemble_poisson(self, gamma=1.0, source=None, boundary=None, reference=None):
        """assemble_poisson with Diffusivity::Constant(gamma); `boundary` maps
        a boundary face to ("FixedValue", value) or "ZeroGradient" (the
        default); `reference` is (cell, value) or None.  Returns (CsrMatrix
        on the device, rhs)."""
        nf = self.nfaces
        kind = np.zeros(nf, np.int32)
        value = np.zeros(nf)
        for f, bc in (boundary or {}).items():
            if bc != "ZeroGradient":
                kind[f], value[f] = FIXED_VALUE, bc[1]
        src = None if source is None else np.ascontiguousarray(source, dtype=np.float64)
        rhs = np.empty(self.ncells)
        h = C.c_void_p()
        pk = C.c_int32()
        detail = C.c_double()
        has_ref = reference is not None
        rc, rv = reference if has_ref else (0, 0.0)
        st = load().fl_mesh2d_assemble_poisson(self.handle, gamma, ptr(src), ptr(kind), ptr(value),
                                               int(has_ref), rc, rv, C.byref(h), ptr(rhs),
                                               C.byref(pk), C.byref(detail))
        if st == 15:
            raise PoissonError(POISSON_KINDS.get(pk.value, str(pk.value)), _last_error(self.ctx),
                               detail.value)
        check(st, self.ctx.handle)
        return CsrMatrix.adopt(h, self.ctx), rhs
