"""Helpers shared by the GPU parity tests: run one DoStep through the C ABI from a
given state, and the velocity tolerance of the parity bar."""
import numpy as np

from umans_amd import gpu

REL_TOL = 1e-5        # relative velocity error bar against the FP64 oracle
FLOOR = 1e-3          # |v| floor of the relative error (velocities can be 0 or clamped)


def gpu_step(world_type, w, h, dt, packed, agents, max_neighbors=0, steps=1, record=True, cell_size=0.0,
             obstacles=None):
    """obstacles: optional list of polygons (vertex lists)."""
    W = gpu.World(world_type, w, h, dt, packed, max_neighbors=max_neighbors, cell_size=cell_size)
    try:
        if obstacles:
            W.set_obstacles([len(p) for p in obstacles], np.array([xy for p in obstacles for xy in p], np.float32))
        W.upload(agents)
        if record:
            W.record_neighbors(True)
        W.step(steps)
        out = W.download()
        nb = W.neighbors() if record else None
        st = W.stats()
    finally:
        W.close()
    return out, nb, st


def rel_err(v, ref):
    d = np.linalg.norm(v.astype(np.float64) - ref.astype(np.float64), axis=1)
    return d / np.maximum(np.linalg.norm(ref.astype(np.float64), axis=1), FLOOR)
