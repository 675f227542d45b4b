"""Discrete adjoint of a run with remeshing: a frozen dt and mesh sequence.

The forward run steps a mesh that is remeshed on the device between steps
(swe_adapt, the reference's adaption traversal, SWE.f90:387-390).  It
records everything the reverse sweep needs:

- per step, the mesh it ran on, its dt, the wet/dry limiter masks and the
  state before it;
- per remeshing, its transfer record (swe_adapt_record_device).

With all of that frozen, the run is a composition of traversals F_k and
state transfers T_k. A transfer is linear: a split copies, a merge
averages. The reverse sweep applies the adjoint traversal of each step on
that step's mesh (swe_adjoint_step_device) and the transpose of each
transfer (swe_transfer_adjoint_device).

The loss is a weighted sum of the surface elevation after each step,
L = sum_k <w_k, eta_{k+1}>, with w_k over the cells of step k's mesh.
(Gauge misfits are of this form, with w_k the misfit's derivative at the
gauges' cells.)  The reverse sweep returns dL/d(eta_0, hu_0, hv_0, b_0).

Every state is kept (no checkpoints). This suits runs whose meshes and
states fit in HBM together.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native, swe


class RemeshedRun:
    """Forward run with remeshing on `mesh` with EulerTimestep parameters
    `params`; torch device tensors hold the tape."""

    def __init__(self, mesh: swe.Mesh, params: dict, device: str = "cuda:0"):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.params = dict(params)
        self.op = swe.EulerTimestep(mesh, **params)
        self.meshes = [(mesh.depth.copy(), mesh.plotter.copy(), np.asarray(mesh.nbr).copy())]
        self.steps = []      # (mesh index, dt, forward, eta, hu, hv, b, masks)
        self.records = {}    # step -> (code, base, b_sampled) of the remeshing before it
        self._ops = {}       # handles on earlier meshes of the sequence, for the reverse sweep

    def _dev(self, a):
        return self.torch.as_tensor(np.ascontiguousarray(a), device=self.device)

    def forward(self, state0, dt, n_steps: int, forward: bool, remesh=None):
        """n_steps traversals from state0 = (eta, hu, hv, b) with the dt
        sequence dt (a number or one per step).  remesh(k, n) returns the
        refinement flags (host int8 over the n current cells) of a remeshing
        before step k, or None.  The direction flips after every traversal
        and every remeshing, as in the reference.  Returns the state after
        each step (host arrays) and the final direction."""
        torch = self.torch
        dts = np.broadcast_to(np.asarray(dt, np.float64), (n_steps,))
        eta, hu, hv, b = (np.asarray(a, np.float64).copy() for a in state0)
        m = len(self.meshes) - 1
        outs = []
        P = lambda t: t.data_ptr()  # noqa: E731
        for k in range(n_steps):
            flags = remesh(len(self.steps), self.op.n) if remesh else None
            if flags is not None:
                self.op.upload(eta, hu, hv, b)
                self.op.adapt(np.asarray(flags, np.int8))
                n_old, sampled = C.c_int64(), C.c_int()
                _native.check("swe_adapt_record_device", self.op._lib.swe_adapt_record_device(
                    self.op._h, None, None, C.byref(n_old), C.byref(sampled), None))
                code = torch.empty(n_old.value, dtype=torch.int8, device=self.device)
                base = torch.empty(n_old.value, dtype=torch.int64, device=self.device)
                _native.check("swe_adapt_record_device", self.op._lib.swe_adapt_record_device(
                    self.op._h, P(code), P(base), C.byref(n_old), None,
                    torch.cuda.current_stream().cuda_stream))
                self.records[len(self.steps)] = (code, base, sampled.value)
                eta, hu, hv, b = self.op.download()
                depth, plotter, nbr = self.op.download_mesh()
                self.meshes.append((depth, plotter, nbr))
                m = len(self.meshes) - 1
                forward = not forward
            q = [self._dev(a) for a in (eta, hu, hv, b)]
            out = [torch.empty_like(q[0]) for _ in range(3)]
            dtd = torch.tensor([dts[k]], dtype=torch.float64, device=self.device)
            masks = torch.empty(q[0].shape[0], dtype=torch.int8, device=self.device)
            self.op.forward_device(*map(P, q), *map(P, out), P(dtd), forward, masks=P(masks),
                                   stream=torch.cuda.current_stream().cuda_stream)
            self.steps.append((m, dtd, forward, *q, masks))
            eta, hu, hv = (t.cpu().numpy() for t in out)
            outs.append((eta.copy(), hu.copy(), hv.copy(), b.copy()))
            forward = not forward
        torch.cuda.synchronize()
        return outs, forward

    def _op_for(self, m: int) -> swe.EulerTimestep:
        """A handle on mesh m of the sequence: the forward handle holds the
        last one; the earlier ones get handles of their own."""
        if m == len(self.meshes) - 1:
            return self.op
        if m not in self._ops:
            depth, plotter, nbr = self.meshes[m]
            self._ops[m] = swe.EulerTimestep(swe.Mesh(depth, plotter, nbr), **self.params)
        return self._ops[m]

    def backward(self, weights):
        """weights[k]: dL/d(eta after step k) (host or device arrays over the
        cells of step k's mesh).  Returns dL/d(eta_0, hu_0, hv_0, b_0) on
        the first mesh, as device tensors."""
        torch = self.torch
        P = lambda t: t.data_ptr()  # noqa: E731
        stream = torch.cuda.current_stream().cuda_stream
        n_steps = len(self.steps)
        le = lu = lv = gb = None
        for k in range(n_steps - 1, -1, -1):
            m, dtd, fwd, e, u, v, b, masks = self.steps[k]
            w = self._dev(weights[k]) if not torch.is_tensor(weights[k]) else weights[k]
            if le is None:
                le, lu, lv = w.clone(), torch.zeros_like(w), torch.zeros_like(w)
                gb = torch.zeros_like(w)
            else:
                le = le + w
            lo = [torch.empty_like(le) for _ in range(3)]
            op = self._op_for(m)
            op.adjoint_device(P(e), P(u), P(v), P(b), P(masks), P(le), P(lu), P(lv),
                              *map(P, lo), P(gb), P(dtd), fwd, stream=stream)
            le, lu, lv = lo
            if k in self.records:
                code, base, sampled = self.records[k]
                n_old = code.shape[0]
                to = [torch.empty(n_old, dtype=torch.float64, device=self.device)
                      for _ in range(3)]
                go = torch.zeros(n_old, dtype=torch.float64, device=self.device)
                _native.check("swe_transfer_adjoint_device",
                              self.op._lib.swe_transfer_adjoint_device(
                                  n_old, P(code), P(base), sampled, P(le), P(lu), P(lv), P(gb),
                                  *map(P, to), P(go), stream))
                le, lu, lv = to
                gb = go
        torch.cuda.synchronize()
        return le, lu, lv, gb

    def close(self):
        for op in self._ops.values():
            op.close()
        self._ops = {}
        self.op.close()
