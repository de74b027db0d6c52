"""The integration scheduler of the reference (pandia_integration,
integrate.cpp) over a GPUVoxelGrid: frames wait in an integration buffer, the
integrated frames whose world pose has moved away from the pose they were
integrated with are selected for reintegration, and removed frames are
deintegrated.

The buffers and their rules follow integrate.cpp:
  * integrate(f, T_world(f)); f.integrateddofs = MattoDof(T_world(f))  (:94-105, :148-160)
  * reintegrate(f): deIntegrate(f, getT(f.integrateddofs)), then
    integrate(f, T_world(f)); f.integrateddofs = worlddofs(f)         (:106-125)
  * deintegrate(f): deIntegrate(f, getT(f.integrateddofs)); f leaves the
    integrated list                                                    (:127-141)
  * updateReintegrationBuffer: the integrated frames sorted by diff, largest
    first, down to the first with diff < treint                       (:211-234)
  * removeFrameFromIntegration: a waiting frame leaves the buffer, an
    integrated one is queued for deintegration                         (:21-44)
Poses reach the grid as the float cast of the double matrix, the
deintegration pose as the float cast of getT(integrateddofs) (which can differ
from the integrated pose in the last float bit, as in the reference).

step() is one pass of the reference's thread loop (:60-141: at most one
integration, one reintegration and one deintegration, in that order, as plain
calls).  run_batched() drains the buffers instead: all waiting integrations,
then the whole reintegration buffer as ONE batched launch (tsdf_reintegrate_batch,
bit-identical to the same plain calls in the same order), then the
deintegrations.  finish() is the stop path (:148-207).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import check

G_TREINT = 0.04  # treint of the reference's config.txt:33


def mat_to_dof(T) -> np.ndarray:
    """MattoDof, matrixutil.cpp:38-64: (rx, ry, rz, tx, ty, tz)."""
    a = np.ascontiguousarray(np.asarray(T, np.float64).reshape(16))
    out = np.zeros(6, np.float64)
    dp = C.POINTER(C.c_double)
    check(_lib.load().tsdf_mat_to_dof(a.ctypes.data_as(dp), out.ctypes.data_as(dp)), "tsdf_mat_to_dof")
    return out


def dof_to_mat(dof) -> np.ndarray:
    """getT, matrixutil.h:61-63: trans(t) Rz Ry Rx."""
    a = np.ascontiguousarray(np.asarray(dof, np.float64).reshape(6))
    out = np.zeros(16, np.float64)
    dp = C.POINTER(C.c_double)
    check(_lib.load().tsdf_dof_to_mat(a.ctypes.data_as(dp), out.ctypes.data_as(dp)), "tsdf_dof_to_mat")
    return out.reshape(4, 4)


def reint_diff(world_dof, integrated_dof) -> float:
    """diff, integrate.cpp:211-216."""
    a = np.ascontiguousarray(np.asarray(world_dof, np.float64).reshape(6))
    b = np.ascontiguousarray(np.asarray(integrated_dof, np.float64).reshape(6))
    dp = C.POINTER(C.c_double)
    return float(_lib.load().tsdf_reint_diff(a.ctypes.data_as(dp), b.ctypes.data_as(dp)))


def reint_select(world_dofs, integrated_dofs, treint=G_TREINT):
    """updateReintegrationBuffer's sort and cut: (list indices by diff,
    largest first, equal ones in list order; how many lead with diff >= treint)."""
    w = np.ascontiguousarray(np.asarray(world_dofs, np.float64).reshape(-1, 6))
    i = np.ascontiguousarray(np.asarray(integrated_dofs, np.float64).reshape(-1, 6))
    n = w.shape[0]
    if i.shape[0] != n:
        raise ValueError("world and integrated dofs of different lengths")
    order = np.zeros(max(n, 1), np.int32)
    m = C.c_int32(0)
    dp = C.POINTER(C.c_double)
    check(_lib.load().tsdf_reint_select(n, w.ctypes.data_as(dp), i.ctypes.data_as(dp), float(treint),
                                        order.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(m)),
          "tsdf_reint_select")
    return [int(k) for k in order[:n]], int(m.value)


@dataclass(eq=False)
class Frame:
    """The parts of the reference's Frame (Frame.h) the scheduler uses:
    device depth (uint16 mm) and optional rgb, the world pose and the dofs the
    frame was integrated with."""
    depth: torch.Tensor
    rgb: torch.Tensor | None = None
    frametoworldtrans: np.ndarray = field(default_factory=lambda: np.eye(4))
    worlddofs: np.ndarray = field(default_factory=lambda: np.zeros(6))
    integrateddofs: np.ndarray = field(default_factory=lambda: np.zeros(6))
    integrated: bool = False
    unique_id: int = -1

    def setFrametoWorldTrans(self, T):
        """Frame.cpp:71-74: the matrix and its dofs."""
        self.frametoworldtrans = np.asarray(T, np.float64).reshape(4, 4).copy()
        self.worlddofs = mat_to_dof(self.frametoworldtrans)


def _f32(T):
    return np.asarray(T, np.float64).astype(np.float32)


class IntegrationScheduler:
    """pandia_integration's buffers over one grid (single-threaded: the
    caller serialises, as the reference does with integrationlock)."""

    def __init__(self, grid, treint: float = G_TREINT, with_color: bool = True):
        self.grid = grid
        self.treint = float(treint)
        self.with_color = bool(with_color) and grid.withColor
        self.integrationBuffer: list[Frame] = []
        self.reintegrationBuffer: list[Frame] = []
        self.deintegrationBuffer: list[Frame] = []
        self.integratedframes: list[Frame] = []

    # -- buffers -----------------------------------------------------------
    def push(self, f: Frame) -> None:
        self.integrationBuffer.append(f)

    def removeFrameFromIntegration(self, f: Frame) -> None:
        """integrate.cpp:21-44."""
        for k, g in enumerate(self.integrationBuffer):
            if g is f:
                del self.integrationBuffer[k]
                return
        if any(g is f for g in self.integratedframes):
            self.deintegrationBuffer.append(f)

    def updateReintegrationBuffer(self) -> None:
        """integrate.cpp:223-234 (the integrated list is left in diff order,
        as list::sort leaves it)."""
        fr = self.integratedframes
        if not fr:
            return
        order, m = reint_select(np.stack([f.worlddofs for f in fr]), np.stack([f.integrateddofs for f in fr]),
                                self.treint)
        self.integratedframes = [fr[k] for k in order]
        self.reintegrationBuffer = self.integratedframes[:m]

    # -- grid calls -----------------------------------------------------------
    def _rgb(self, f):
        return f.rgb if self.with_color else None

    def _integrate(self, f, T, sign):
        call = self.grid.integrate if sign > 0 else self.grid.deIntegrate
        call(self._rgb(f), f.depth, _f32(T), withColor=self.with_color)

    def _do_integrate(self, f: Frame) -> None:
        self.integratedframes.append(f)
        self._integrate(f, f.frametoworldtrans, +1)
        f.integrateddofs = f.worlddofs.copy()
        f.integrated = True

    def _do_reintegrate(self, f: Frame) -> None:
        self._integrate(f, dof_to_mat(f.integrateddofs), -1)
        self._integrate(f, f.frametoworldtrans, +1)
        f.integrateddofs = f.worlddofs.copy()

    def _do_deintegrate(self, f: Frame) -> None:
        self.integratedframes = [g for g in self.integratedframes if g is not f]
        self._integrate(f, dof_to_mat(f.integrateddofs), -1)
        f.integrateddofs = np.zeros(6)

    def step(self) -> bool:
        """One pass of the thread loop, integrate.cpp:60-141; False when
        every buffer was empty."""
        fint = self.integrationBuffer.pop(0) if self.integrationBuffer else None
        fre = self.reintegrationBuffer.pop(0) if self.reintegrationBuffer else None
        fde = self.deintegrationBuffer.pop(0) if self.deintegrationBuffer else None
        if fde is not None:
            self.integratedframes = [g for g in self.integratedframes if g is not fde]
        if fint is not None:
            self._do_integrate(fint)
        if fre is not None:
            self._do_reintegrate(fre)
        if fde is not None:
            self._integrate(fde, dof_to_mat(fde.integrateddofs), -1)
            fde.integrateddofs = np.zeros(6)
        return not (fint is None and fre is None and fde is None)

    def run_batched(self) -> dict:
        """Drains the buffers: the waiting integrations (plain calls), the
        reintegration buffer as one batched launch (deIntegrate(getT(dofs)),
        integrate(T_world) per frame, in buffer order), the deintegrations.
        Returns the counts."""
        n_int = len(self.integrationBuffer)
        while self.integrationBuffer:
            self._do_integrate(self.integrationBuffer.pop(0))
        re = self.reintegrationBuffer
        self.reintegrationBuffer = []
        if re:
            old = [_f32(dof_to_mat(f.integrateddofs)) for f in re]
            new = [_f32(f.frametoworldtrans) for f in re]
            rgbs = [f.rgb for f in re] if self.with_color else None
            self.grid.reintegrate_batch([f.depth for f in re], old, new, rgbs=rgbs)
            for f in re:
                f.integrateddofs = f.worlddofs.copy()
        n_de = len(self.deintegrationBuffer)
        while self.deintegrationBuffer:
            self._do_deintegrate(self.deintegrationBuffer.pop(0))
        return {"integrated": n_int, "reintegrated": len(re), "deintegrated": n_de}

    def finish(self) -> None:
        """The stop path, integrate.cpp:148-207.  Its last loop integrates
        (not deintegrates) the frames still waiting for deintegration with
        getT(integrateddofs) (:196-207); that is restated as written."""
        while self.integrationBuffer:
            f = self.integrationBuffer.pop(0)
            self._integrate(f, f.frametoworldtrans, +1)
            f.integrateddofs = f.worlddofs.copy()
            f.integrated = True
            self.integratedframes.append(f)
        while self.reintegrationBuffer:
            f = self.reintegrationBuffer.pop(0)
            self._integrate(f, dof_to_mat(f.integrateddofs), -1)
            self._integrate(f, f.frametoworldtrans, +1)
            f.integrateddofs = f.worlddofs.copy()
        while self.deintegrationBuffer:
            f = self.deintegrationBuffer.pop(0)
            self._integrate(f, dof_to_mat(f.integrateddofs), +1)
            f.integrateddofs = f.worlddofs.copy()
            self.integratedframes = [g for g in self.integratedframes if g is not f]
