"""The adjoint of a run with remeshing (a frozen dt and mesh sequence,
samoa_amd/adjoint_adaptive.py) against torch.autograd of the same run
restated in PyTorch: the FP64 traversal of tests/torch_ref.py on each
step's mesh and the state transfers of the remeshings (a split copies, a
merge averages) built from the recorded transfer codes.  The forward run is
bit-identical to the restatement; the gradients agree to rounding,
max |x - y| <= 1e-9 max |y| per field."""
import numpy as np
import pytest
import torch

import pyoracle as po
import torch_ref as tr
from samoa_amd import host, swe
from samoa_amd.adjoint_adaptive import RemeshedRun
from test_gpu_adjoint import wetdry_state

pytestmark = pytest.mark.gpu

SPLIT_CELLS = {4: 2, 5: 3, 6: 3, 7: 4}
NORMALS = None


def torch_mesh(depth, plotter, nbr):
    global NORMALS
    if NORMALS is None:
        NORMALS = np.stack([po.normals(p) if p else np.zeros((3, 2)) for p in range(-8, 9)])
    return tr.TorchMesh(depth, plotter, nbr, NORMALS, po.leg_sizes())


def transfer_torch(q, code, base):
    """The state transfer of a remeshing record, as torch index operations."""
    code = code.astype(np.int64)
    b0 = base & 0xffffffff
    n_new = int(sum(1 if k in (0, 1) else (0 if k == 2 else SPLIT_CELLS[k]) for k in code))
    src1 = np.zeros(n_new, np.int64)
    src2 = np.zeros(n_new, np.int64)
    merged = np.zeros(n_new, bool)
    for c, (k, i) in enumerate(zip(code, b0)):
        if k == 0:
            src1[i] = src2[i] = c
        elif k == 1:
            src1[i], src2[i], merged[i] = c, c + 1, True
        elif k >= 4:
            src1[i:i + SPLIT_CELLS[k]] = c
            src2[i:i + SPLIT_CELLS[k]] = c
    s1, s2, m = torch.tensor(src1), torch.tensor(src2), torch.tensor(merged)
    return [torch.where(m, 0.5 * (a[s1] + a[s2]), a[s1]) for a in q]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_adjoint_through_remeshing_matches_autograd(seed):
    d, S, drytol, n_steps, dt = 7, 10.0, 1e-3, 7, 2e-3
    mesh = host.mesh_uniform(d)
    st = wetdry_state(mesh, seed, S)
    params = dict(scaling=S, dry_tolerance=drytol, min_depth=d - 2, max_depth=d + 2,
                  boundary=swe.BC_WALL)
    rng = np.random.default_rng(seed)

    def remesh(k, n):  # random requests before steps 2 and 4; nothing changes before 5
        if k in (2, 4):
            return rng.choice(np.array([-1, 0, 1], np.int8), size=n, p=[0.4, 0.3, 0.3])
        if k == 5:
            return np.zeros(n, np.int8)
        return None

    run = RemeshedRun(mesh, params)
    outs, _ = run.forward(st, dt, n_steps, False, remesh=remesh)
    sizes = [len(m[0]) for m in run.meshes]
    print("case: seed", seed, "meshes", sizes, flush=True)
    assert len(sizes) == 4 and sizes[1] != sizes[0] and sizes[3] == sizes[2]
    weights = [rng.normal(size=len(o[0])) for o in outs]
    g_dev = [t.cpu().numpy() for t in run.backward(weights)]

    # the same run in torch, with autograd
    q0 = [torch.tensor(a, requires_grad=True) for a in st]
    q = list(q0)
    loss = 0.0
    fwd = False
    for k, (m, dtd, f, *_rest) in enumerate(run.steps):
        if k in run.records:
            code, base, _ = run.records[k]
            q = transfer_torch(q, code.cpu().numpy(), base.cpu().numpy())
            fwd = not fwd
        assert f == fwd
        depth, plotter, nbr = run.meshes[m]
        e, u, v, _ = tr.step(torch_mesh(depth, plotter, nbr), *q, dt, fwd, S, po.G, drytol,
                             bc=1)
        assert np.array_equal(e.detach().numpy(), outs[k][0]), f"step {k}"
        assert np.array_equal(u.detach().numpy(), outs[k][1]), f"step {k}"
        loss = loss + (e * torch.tensor(weights[k])).sum()
        q = [e, u, v, q[3]]
        fwd = not fwd
    g_ref = [g.numpy() for g in torch.autograd.grad(loss, q0)]
    run.close()
    for name, x, y in zip(("eta", "hu", "hv", "b"), g_dev, g_ref):
        assert np.abs(x - y).max() <= 1e-9 * max(np.abs(y).max(), 1e-300), name
