"""Frame output and restart files in the reference's conventions (host logic)."""
import os
import pickle

import numpy as np
import pytest

from fastipc_amd import io as fio
from fastipc_amd import scene as S


def test_obj_matches_reference_format(tmp_path):
    sc = S.tiny_pair()
    p, e, t = sc.boundary
    path = str(tmp_path / "000001.obj")
    fio.write_obj(path, sc.x, t)
    lines = open(path).read().splitlines()
    assert len(lines) == len(sc.x) + len(t)
    # NEWTON_IPC.py:927-930: 'v %.6f %.6f %.6f' then 1-based 'f %d %d %d'
    assert lines[0] == "v %.6f %.6f %.6f" % tuple(sc.x[0])
    assert lines[len(sc.x)] == "f %d %d %d" % tuple(t[0] + 1)
    x, f = fio.read_obj(path)
    assert np.abs(x - sc.x).max() <= 5e-7 and np.array_equal(f, t)


def test_cache_round_trip_and_reference_layout(tmp_path):
    rng = np.random.default_rng(0)
    x, v = rng.normal(size=(50, 3)), rng.normal(size=(50, 3))
    w = fio.FrameWriter(str(tmp_path), np.array([[0, 1, 2]]))
    w.write(7, x, v)
    assert os.path.exists(tmp_path / "objs" / "000007.obj")
    cp = tmp_path / "caches" / "000007.p"
    x2, v2 = fio.load_cache(str(cp))
    assert np.array_equal(x, x2) and np.array_equal(v, v2)
    # what the reference's restart reads: a pickled [x, v] list
    with open(cp, "rb") as f:
        raw = f.read()
    assert raw[:1] == b"\x80"


def test_cache_reader_refuses_other_objects(tmp_path):
    p = tmp_path / "bad.p"
    with open(p, "wb") as f:
        pickle.dump([np.zeros(3), os.getcwd], f)
    with pytest.raises(pickle.UnpicklingError):
        fio.load_cache(str(p))


def test_npz_checkpoint(tmp_path):
    x, v = np.ones((4, 3)), np.zeros((4, 3))
    p = str(tmp_path / "c.npz")
    fio.save_checkpoint(p, x, v, 12)
    x2, v2, f = fio.load_checkpoint(p)
    assert np.array_equal(x, x2) and np.array_equal(v, v2) and f == 12
