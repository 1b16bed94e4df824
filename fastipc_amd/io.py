"""Frame output and restart files in the reference's conventions.

* OBJ frames: ``objs/{f:06d}.obj`` with ``v %.6f %.6f %.6f`` per vertex and
  ``f a b c`` (1-based) per boundary triangle, as
  projects/FastIPC/NEWTON_IPC.py:925-930 writes them.
* Restart caches: ``caches/{f:06d}.p``, a pickled list ``[x, v]`` of float64
  arrays (NEWTON_IPC.py:1002, README.md:65-75), so that either code can restart
  from the other's frames.  Reading uses an unpickler that rebuilds only NumPy
  arrays and lists; anything else in the file is refused.
* ``save_checkpoint`` / ``load_checkpoint``: the same state as ``.npz``.
"""
from __future__ import annotations

import io as _io
import os
import pickle

import numpy as np


def write_obj(path: str, x: np.ndarray, triangles: np.ndarray) -> None:
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(triangles, dtype=np.int64) + 1
    with open(path, "w") as f:
        f.write("".join("v %.6f %.6f %.6f\n" % tuple(p) for p in x))
        f.write("".join("f %d %d %d\n" % tuple(tri) for tri in t))


def read_obj(path: str):
    """(vertices, triangles 0-based) of an OBJ written by write_obj."""
    vs, fs = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                vs.append([float(a) for a in line.split()[1:4]])
            elif line.startswith("f "):
                fs.append([int(a.split("/")[0]) - 1 for a in line.split()[1:4]])
    return np.array(vs, np.float64), np.array(fs, np.int64)


class _ArrayUnpickler(pickle.Unpickler):
    """Rebuilds only NumPy arrays (and the containers of a [x, v] list)."""
    ALLOWED = {
        ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
        ("numpy", "ndarray"), ("numpy", "dtype"), ("numpy.core.multiarray", "scalar"),
        ("numpy._core.multiarray", "scalar"),
    }

    def find_class(self, module, name):
        if (module, name) in self.ALLOWED:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"refusing {module}.{name} in a restart cache")


def save_cache(path: str, x: np.ndarray, v: np.ndarray) -> None:
    with open(path, "wb") as f:
        pickle.dump([np.asarray(x, np.float64), np.asarray(v, np.float64)], f)


def load_cache(path: str):
    with open(path, "rb") as f:
        data = _ArrayUnpickler(_io.BytesIO(f.read())).load()
    x, v = data[0], data[1]
    return np.asarray(x, np.float64), np.asarray(v, np.float64)


def save_checkpoint(path: str, x: np.ndarray, v: np.ndarray, frame: int) -> None:
    np.savez(path, x=np.asarray(x, np.float64), v=np.asarray(v, np.float64), frame=np.int64(frame))


def load_checkpoint(path: str):
    d = np.load(path, allow_pickle=False)
    return d["x"], d["v"], int(d["frame"])


class FrameWriter:
    """Per-frame OBJ and restart cache in the reference's directory layout."""

    def __init__(self, directory: str, triangles: np.ndarray, caches: bool = True):
        self.dir = directory
        self.tris = np.asarray(triangles)
        self.caches = caches
        os.makedirs(os.path.join(directory, "objs"), exist_ok=True)
        if caches:
            os.makedirs(os.path.join(directory, "caches"), exist_ok=True)

    def write(self, f: int, x: np.ndarray, v: np.ndarray | None = None) -> None:
        write_obj(os.path.join(self.dir, "objs", f"{f:06d}.obj"), x, self.tris)
        if self.caches and v is not None:
            save_cache(os.path.join(self.dir, "caches", f"{f:06d}.p"), x, v)


def restart(sim, path: str) -> None:
    """Set a NewtonIPC state from a restart cache (or an .npz checkpoint)."""
    if path.endswith(".npz"):
        x, v, frame = load_checkpoint(path)
        sim.frame = frame
    else:
        x, v = load_cache(path)
        sim.frame = int(os.path.basename(path).split(".")[0])
    sim.set_state(x, v)
