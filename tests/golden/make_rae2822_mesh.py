"""Writes tests/golden/rae2822_polymesh.npz: the RAE2822 tutorial's polyMesh
(tutorials/RAE2822/constant/polyMesh of the reference, a data file of its
tutorial) as arrays, so that the tests run the case where no reference
checkout is present.  Read with the package's own ASCII polyMesh reader; run
where the reference tree is:

    python tests/golden/make_rae2822_mesh.py /root/reference/tutorials/RAE2822
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from rhodst_amd import meshgen  # noqa: E402


def main(case_dir: str) -> None:
    m = meshgen.read_polymesh(case_dir)
    names = np.array([p[0] for p in m.patches])
    meta = np.array([[p[1], p[2], p[3]] for p in m.patches], dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "rae2822_polymesh.npz"), points=m.points, face_offsets=m.face_offsets,
                        face_points=m.face_points, owner=m.owner, neighbour=m.neighbour, n_cells=np.int64(m.n_cells),
                        patch_names=names, patch_meta=meta)
    print(f"{m.n_cells} cells, {len(m.owner)} faces, {len(m.neighbour)} internal, patches {m.patches}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tutorials/RAE2822")
