"""Many small problems advanced in one persistent, graph-captured launch
(tbx_batch): every problem is a velocity tree and a one-component
concentration tree of degree 6, 8 or 10 with its own time step, stepped in
FP32 with the refit fused per leaf.  The launch configuration (workgroup
size x leaves per work item) is chosen by timing each on the batch."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BC_FREESPACE, check, lib
from .tree import Tree, _stream

CONFIGS = [(256, 1), (256, 2), (512, 1), (512, 2), (512, 4), (1024, 1)]


class Batch:
    def __init__(self, vel, conc, dt, bc: int = BC_FREESPACE):
        if not (len(vel) == len(conc) == len(dt)) or not vel:
            raise ValueError("one velocity tree, concentration tree and dt per problem")
        self.vel, self.conc = list(vel), list(conc)  # keep the trees alive
        self._attached = []
        n = len(vel)
        vh = (C.c_void_p * n)(*[t.handle.value for t in vel])
        ch = (C.c_void_p * n)(*[t.handle.value for t in conc])
        dts = np.ascontiguousarray(dt, dtype=np.float64)
        h = C.c_void_p()
        check(lib().tbx_batch_create(n, vh, ch, dts.ctypes.data, bc, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().tbx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def config(self) -> int:
        c = C.c_int()
        check(lib().tbx_batch_get_config(self._h, C.byref(c)))
        return c.value

    @property
    def node_velocity(self) -> bool:
        """Whether the first RK2 stage reads the velocities' nodal values
        (decided when the batch was made)."""
        c = C.c_int()
        check(lib().tbx_batch_node_velocity(self._h, C.byref(c)))
        return bool(c.value)

    def set_config(self, cfg: int) -> None:
        check(lib().tbx_batch_set_config(self._h, cfg))

    def tune(self):
        """Time every configuration once; keep the fastest.  Returns ms per
        configuration."""
        ms = np.zeros(len(CONFIGS))
        check(lib().tbx_batch_tune(self._h, ms.ctypes.data, _stream()))
        return ms

    def attach_step(self, vel: Tree, conc: Tree, dt: float, num_rk_step: int = 1,
                    mode: str = "fp32") -> None:
        """Carry an in-situ step of (vel, conc) (stationary velocity) in every
        replay of the batch's step graph, ahead of the persistent kernel."""
        from .tree import _mode

        self._attached.append((vel, conc))  # keep the trees alive
        check(lib().tbx_batch_attach_step(self._h, vel.handle, conc.handle, dt, num_rk_step,
                                          _mode(mode)))

    def step(self, nsteps: int = 1) -> None:
        check(lib().tbx_batch_step(self._h, nsteps, _stream()))

    def coeffs(self, p: int) -> np.ndarray:
        t = self.conc[p]
        out = np.empty((t.L, 1, t.nc))
        check(lib().tbx_batch_get_coeffs(self._h, p, out.ctypes.data, _stream()))
        return out
