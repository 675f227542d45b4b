"""Loaders for the committed golden fixtures (tests/golden/*.npz)."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def runs():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN, "run_*.npz"))):
        z = np.load(f, allow_pickle=False)
        meta = json.loads(str(z["meta"]))
        out.append((os.path.basename(f)[4:-4], z, meta))
    return out


def direction_schedule(steps, nout):
    """Direction of each Euler traversal of a reference run with -nout nout.

    Traversal parity: init_grid ends reversed (SFC_node_traversal.f90:82), so
    init_b runs backward, init_dofs forward, the initial output backward
    (SWE.f90:230-297); every traversal flips the direction
    (SFC_generic_traversal_ringbuffer.f90:420)."""
    trav = 3
    fwd = []
    for k in range(1, steps + 1):
        trav += 1
        fwd.append(trav % 2 == 0)
        if k % nout == 0:
            trav += 1
    return fwd
