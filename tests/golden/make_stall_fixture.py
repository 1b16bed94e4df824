"""Writes tests/golden/stall_local.npz: the local geometry where the
reference's Newton loop stalls on a 3x3x3 grid of 5^3-vertex cubes with the
reference's absolute Newton tolerance (1e-2 dt).

    python tests/golden/make_stall_fixture.py      (about two minutes)

The oracle (exact direct solve) runs 45 steps of cube_grid(5, 3, seed=0);
in step 46 the line search stops making progress after 15 Newton iterations.
The fixture keeps the six vertices of the two boundary triangles sharing the
edge (1295, 1320) and of the edge (1651, 1675) at the stalled iterate, the
Newton direction there, their rest positions, dHat, kappa and the hash cell
size."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fastipc_amd import scene as S  # noqa: E402
from oracle.newton import OracleIPC  # noqa: E402

V = [1290, 1295, 1296, 1320, 1651, 1675]
TRIS = [[1295, 1296, 1320], [1295, 1320, 1290]]
EDGES = [[1295, 1320], [1651, 1675]]


def main():
    sc = S.cube_grid(5, 3, seed=0)
    o = OracleIPC(sc)
    for _ in range(45):
        o.step()
    f = o.frame
    o.compute_xn_and_xTilde(f * o.dt)
    # the Newton iterations of step 46 up to the first that cannot lower E
    for it in range(40):
        o.find_constraints()
        o.assemble()
        o.solve_system(f * o.dt)
        E0 = o.compute_energy()
        o.save_xPrev()
        a = o.compute_intersection_free_step_size()
        o.apply_sol(a)
        while o.check_collision():
            a /= 2.0
            o.apply_sol(a)
        o.find_constraints()
        E = o.compute_energy()
        while E > E0 and a > 1e-12:
            a *= 0.5
            o.apply_sol(a)
            o.find_constraints()
            E = o.compute_energy()
        if a < 1e-6:
            break
    loc = {v: i for i, v in enumerate(V)}
    np.savez(os.path.join(HERE, "stall_local.npz"), x=o.xprev[V], p=o.p[V], x0=o.x0[V],
             tris=np.array([[loc[v] for v in t] for t in TRIS], np.int32),
             edges=np.array([[loc[v] for v in e] for e in EDGES], np.int32),
             dhat=o.dHat, kappa=o.kappa, inv_dx=o.inv_dx, newton_iteration=it)
    print(f"wrote stall_local.npz (stalled at Newton iteration {it} of step {f + 1})")


if __name__ == "__main__":
    main()
