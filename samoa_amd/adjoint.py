"""Gradients of gauge time series through the SWE time loop (discrete adjoint).

`GaugeInversion` freezes one forward run of the scheme: the mesh (or a batch
of meshes advanced by one launch per step), the traversal directions
(alternating, as the reference traverses, SWE.f90:393-403) and the dt
sequence of that run (dt_{n+1} = courant * CFL minimum of step n, as
SWE_euler_timestep.f90:340-343 and SWE.f90:395-398 chain it).
`gauge_series(eta0, b)` then maps the initial surface elevation and the
bathymetry (FP64 device tensors; concatenated over the scenarios of a batch)
to the surface elevation at the gauge cells after every step, as a
differentiable PyTorch operation:

    G = problem.gauge_series(eta0, b)          # (n_steps, n_gauges)
    loss = ((G - target) ** 2).sum()
    loss.backward()                            # eta0.grad, b.grad

The backward pass is the hand-written discrete adjoint
(swe_adjoint_step_device): with dt and the wet/dry limiter decisions of the
forward pass frozen, lam^n = (dF/dQ^n)^T lam^{n+1} + dL/dQ^n and
gb += (dF/db)^T lam^{n+1}.  The forward sweep is taped in one of two ways
(plan_tape): when every state and mask fits the HBM budget (a depth-20 mesh
needs 50 MB per step, so 200 steps take 10 GB of the 288 GB), all of them
are kept and the backward sweep runs the adjoint kernels only; otherwise
memory is bounded by two-level checkpointing: the forward pass keeps every
K-th state (K = ceil(sqrt(n_steps)), so about 3 sqrt(n_steps) states are
alive at once) and the backward pass recomputes one segment of K states at a
time, recording the limiter masks, then runs the adjoint through it
backwards.  The recompute costs one extra forward sweep; it runs on a second
stream, one segment ahead of the adjoint, so the two share the GPU.

`PooledGaugeInversion` runs independent instances (one mesh each) with a
checkpoint pool per instance (plan_pools): store-all or sqrt(N) chosen per
instance against the HBM budget, and admission rounds when even the
checkpointed tapes of all instances exceed it.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .swe import Batch, EulerTimestep, Mesh


def _p(t: torch.Tensor) -> int:
    return t.data_ptr()


class GaugeInversion:
    """A frozen forward problem: mesh(es), gauges, step count, dt sequence.

    `mesh` is a Mesh (with the EulerTimestep keyword arguments in `params`)
    or a list of Meshes with a list of such dicts (one scenario each).
    Gauges index cells of the concatenated state."""

    def __init__(self, mesh, gauges, n_steps: int, params=None, *, forward0: bool = False,
                 checkpoint_every="auto", hbm_budget: float | None = None, device="cuda:0"):
        if isinstance(mesh, Mesh):
            params = dict(params or {})
            self.op = EulerTimestep(mesh, **params)
            self.n = mesh.n
            self.n_scen = 1
            courant = [params.get("courant", 0.45)]
        else:
            params = list(params)
            self.op = Batch(mesh, params)
            self.n = self.op.cells
            self.n_scen = self.op.n_scen
            courant = [p.get("courant", 0.45) for p in params]
        self.device = torch.device(device)
        self.courant = torch.tensor(courant, dtype=torch.float64, device=self.device)
        self.n_steps = int(n_steps)
        if self.n_steps < 1:
            raise ValueError("n_steps must be positive")
        gauges = np.asarray(gauges, np.int64)
        if gauges.ndim != 1 or (gauges < 0).any() or (gauges >= self.n).any():
            raise ValueError("gauges must be cell indices in [0, n)")
        self.gauges = torch.as_tensor(gauges, device=self.device)
        self.forward0 = bool(forward0)
        self.store_all, self.K = self.plan_tape(self.n, self.n_steps, checkpoint_every, hbm_budget,
                                                self.device)
        self.dt = None  # frozen dt sequence, device [n_steps + 1, n_scen]

    @staticmethod
    def plan_tape(n_cells: int, n_steps: int, checkpoint_every="auto", hbm_budget=None,
                  device="cuda:0"):
        """Chooses how the forward sweep is taped.  Returns (store_all, K).

        store_all: every state Q^n (24 B/cell) and the limiter masks of every
        step (1 B/cell) are kept, so the backward sweep runs the adjoint
        kernels only.  Otherwise every K-th state is kept and the backward
        sweep recomputes one segment of K states at a time (two-level
        checkpointing, peak about (n_steps / K + 2 K) states).

        checkpoint_every: "all", an int K, or "auto": store everything when
        it fits in hbm_budget bytes (default: half of the free device memory),
        else K = ceil(sqrt(n_steps))."""
        K_sqrt = max(1, math.ceil(math.sqrt(n_steps)))
        if checkpoint_every == "all":
            return True, 1
        if checkpoint_every != "auto":
            K = int(checkpoint_every)
            if K < 1:
                raise ValueError("checkpoint_every must be >= 1")
            return False, K
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(torch.device(device))
            hbm_budget = 0.5 * free
        state = 24.0 * n_cells
        if n_steps * (state + n_cells) <= hbm_budget:
            return True, 1
        if tape_bytes(n_cells, n_steps, False, K_sqrt) > hbm_budget:
            raise MemoryError(f"{n_cells} cells x {n_steps} steps do not fit a "
                              f"{hbm_budget / 2**30:.1f} GiB checkpoint budget")
        return False, K_sqrt

    def direction(self, step: int) -> bool:
        """Traversal direction of step `step` (0-based): alternating."""
        return self.forward0 ^ (step % 2 == 1)

    def _zeros(self):
        return torch.zeros(self.n, dtype=torch.float64, device=self.device)

    def _empty(self):
        return torch.empty(self.n, dtype=torch.float64, device=self.device)

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _dt(self, step):
        return _p(self.dt) + 8 * self.n_scen * step

    def _check(self, t, name):
        if t.dtype != torch.float64 or t.device != self.device or t.numel() != self.n:
            raise ValueError(f"{name} must be a float64 tensor of {self.n} cells on {self.device}")

    def _fwd(self, q, out, b, step, masks=None, cfl=None, stream=None):
        self.op.forward_device(_p(q[0]), _p(q[1]), _p(q[2]), _p(b), _p(out[0]), _p(out[1]),
                               _p(out[2]), self._dt(step), self.direction(step),
                               dt_cfl=_p(cfl[step]) if cfl is not None else 0,
                               masks=_p(masks) if masks is not None else 0,
                               stream=stream.cuda_stream if stream is not None else self._stream())

    def record_dt(self, eta0: torch.Tensor, b: torch.Tensor, dt0) -> torch.Tensor:
        """Runs the scheme once from (eta0, p = 0, b) with the reference's dt
        chain evaluated on the device and freezes the dt sequence it
        produces; returns it as [n_steps, n_scen]."""
        self._check(eta0, "eta0")
        self._check(b, "b")
        self.dt = torch.empty(self.n_steps + 1, self.n_scen, dtype=torch.float64, device=self.device)
        self.dt[0] = torch.as_tensor(np.broadcast_to(np.asarray(dt0, np.float64), (self.n_scen,)).copy())
        cfl = torch.empty(self.n_steps, self.n_scen, dtype=torch.float64, device=self.device)
        b = b.detach().contiguous()
        q = [eta0.detach().contiguous().clone(), self._zeros(), self._zeros()]
        out = [self._zeros() for _ in range(3)]
        for n in range(self.n_steps):
            self._fwd(q, out, b, n, cfl=cfl)
            torch.mul(self.courant, cfl[n], out=self.dt[n + 1])
            q, out = out, q
        torch.cuda.synchronize(self.device)
        return self.dt[: self.n_steps]

    def gauge_series(self, eta0: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        if self.dt is None:
            raise RuntimeError("record_dt() first: the dt sequence must be frozen")
        self._check(eta0, "eta0")
        self._check(b, "b")
        return _GaugeSeries.apply(eta0, b, self)

    # -- used by the autograd function ------------------------------------
    def run_forward(self, eta0, b):
        """Forward sweep: gauge values after each step and the tape: every
        state and mask (store_all) or the checkpoints Q^{jK}."""
        q = [eta0.detach().contiguous().clone(), self._zeros(), self._zeros()]
        series = torch.empty(self.n_steps, len(self.gauges), dtype=torch.float64, device=self.device)
        if self.store_all:
            # every slot is written by a traversal before it is read
            states = [q] + [[self._empty() for _ in range(3)] for _ in range(self.n_steps - 1)]
            masks = torch.empty(self.n_steps, self.n, dtype=torch.int8, device=self.device)
            last = [self._empty() for _ in range(3)]
            for n in range(self.n_steps):
                out = states[n + 1] if n + 1 < self.n_steps else last
                self._fwd(states[n], out, b, n, masks=masks[n])
                torch.index_select(out[0], 0, self.gauges, out=series[n])
            return series, {"states": states, "masks": masks}
        out = [self._empty() for _ in range(3)]
        checkpoints = {}
        for n in range(self.n_steps):
            if n % self.K == 0:
                checkpoints[n] = [x.clone() for x in q]
            self._fwd(q, out, b, n)
            torch.index_select(out[0], 0, self.gauges, out=series[n])
            q, out = out, q
        return series, {"checkpoints": checkpoints}

    def _adjoint(self, q, mask, b, lam, lam_out, gb, n):
        self.op.adjoint_device(
            _p(q[0]), _p(q[1]), _p(q[2]), _p(b), _p(mask), _p(lam[0]), _p(lam[1]), _p(lam[2]),
            _p(lam_out[0]), _p(lam_out[1]), _p(lam_out[2]), _p(gb), self._dt(n),
            self.direction(n), stream=self._stream())

    def run_backward(self, b, tape, g_series):
        """Adjoint sweep; returns (d/d eta0, d/d b)."""
        lam = [self._zeros() for _ in range(3)]
        lam_out = [self._zeros() for _ in range(3)]
        gb = self._zeros()
        lam[0].index_add_(0, self.gauges, g_series[self.n_steps - 1])
        if self.store_all:
            for n in reversed(range(self.n_steps)):
                self._adjoint(tape["states"][n], tape["masks"][n], b, lam, lam_out, gb, n)
                lam, lam_out = lam_out, lam
                if n >= 1:
                    lam[0].index_add_(0, self.gauges, g_series[n - 1])
            return lam[0].clone(), gb
        # Two-level checkpointing, pipelined: the recompute of segment j - 1
        # runs on a side stream while the adjoint sweeps segment j on the
        # current stream, in the other of two segment buffers.  The two
        # kernels then share the GPU (the adjoint kernel alone holds two
        # waves per SIMD).  Events order every buffer's reuse.
        checkpoints = tape["checkpoints"]
        starts = sorted(checkpoints)
        main = torch.cuda.current_stream(self.device)
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(main)  # the tape and the seeds are written on main
        bufs = []
        for _ in range(2):
            states = [[self._empty() for _ in range(3)] for _ in range(self.K)]
            masks = [torch.empty(self.n, dtype=torch.int8, device=self.device)
                     for _ in range(self.K)]
            spare = [self._empty() for _ in range(3)]
            for t in [x for st in states for x in st] + masks + spare:
                t.record_stream(side)
            bufs.append((states, masks, spare))

        def segment(j):
            s0 = starts[j]
            return s0, (starts[j + 1] if j + 1 < len(starts) else self.n_steps)

        def recompute(j):
            """Q^{s0..s1-1} of segment j and the masks of its steps, on side."""
            states, masks, spare = bufs[j % 2]
            s0, s1 = segment(j)
            with torch.cuda.stream(side):
                for x, y in zip(states[0], checkpoints[s0]):
                    x.copy_(y)
                for n in range(s0, s1):
                    i = n - s0
                    dst = states[i + 1] if i + 1 < s1 - s0 else spare
                    self._fwd(states[i], dst, b, n, masks=masks[i], stream=side)
            ev = torch.cuda.Event()
            ev.record(side)
            return ev

        last = len(starts) - 1
        ready = {last: recompute(last)}
        freed = [None, None]  # the adjoint that last read each buffer
        for j in reversed(range(len(starts))):
            if j >= 1:
                if freed[(j - 1) % 2] is not None:
                    side.wait_event(freed[(j - 1) % 2])
                ready[j - 1] = recompute(j - 1)
            main.wait_event(ready.pop(j))
            states, masks, _ = bufs[j % 2]
            s0, s1 = segment(j)
            for n in reversed(range(s0, s1)):
                self._adjoint(states[n - s0], masks[n - s0], b, lam, lam_out, gb, n)
                lam, lam_out = lam_out, lam
                if n >= 1:
                    lam[0].index_add_(0, self.gauges, g_series[n - 1])
            freed[j % 2] = torch.cuda.Event()
            freed[j % 2].record(main)
        main.wait_stream(side)
        return lam[0].clone(), gb

    def close(self):
        self.op.close()


TAPE_STATE_BYTES = 24   # eta, hu, hv per cell
TAPE_MASK_BYTES = 1     # limiter masks per cell


def tape_bytes(n_cells: int, n_steps: int, store_all: bool, K: int = 1) -> float:
    """HBM a tape takes: every state and mask (store_all), or the
    checkpoints plus two recomputed segments of K states and masks."""
    per = (TAPE_STATE_BYTES + TAPE_MASK_BYTES) * float(n_cells)
    if store_all:
        return n_steps * per
    # the checkpoints, two segment buffers (the backward sweep recomputes one
    # while it runs the adjoint through the other) and their spares
    return (math.ceil(n_steps / K) + 2 * K + 4) * per


def plan_pools(cells, n_steps: int, hbm_budget: float, checkpoint_every="auto"):
    """Per-instance checkpoint pools for independent instances of `cells[s]`
    cells each, admitted against an HBM budget for their tapes.

    Returns a list of rounds; a round is a list of groups (store_all, K,
    [instances]) whose tapes are alive together; rounds run one after the
    other, each freeing its tapes before the next is admitted.

    "auto": every instance keeps all its states if the sum fits; otherwise
    the instances that save the most (the largest) switch to sqrt(N)
    checkpoints (K = ceil(sqrt(n_steps))) until it fits; if even all of them
    checkpointed exceed the budget, instances are admitted in rounds (first
    fit, largest first), and inside a round the smallest instances keep all
    their states while that fits.  An instance whose checkpointed tape alone
    exceeds the budget raises MemoryError.  "all" or an int K: one round with
    that mode for every instance."""
    cells = [int(c) for c in cells]
    idx = list(range(len(cells)))
    K = max(1, math.ceil(math.sqrt(n_steps)))
    if checkpoint_every == "all":
        return [[(True, 1, idx)]]
    if checkpoint_every != "auto":
        Kc = int(checkpoint_every)
        if Kc < 1:
            raise ValueError("checkpoint_every must be >= 1")
        return [[(False, Kc, idx)]]
    full = {s: tape_bytes(cells[s], n_steps, True) for s in idx}
    ckpt = {s: tape_bytes(cells[s], n_steps, False, K) for s in idx}
    for s in idx:
        if ckpt[s] > hbm_budget:
            raise MemoryError(f"instance {s}: {cells[s]} cells x {n_steps} steps do not fit a "
                              f"{hbm_budget / 2**30:.2f} GiB checkpoint budget")

    def modes(members):
        """store-all for as many as fit, smallest first; the rest checkpointed."""
        keep = set()
        total = sum(ckpt[s] for s in members)
        for s in sorted(members, key=lambda s: (full[s] - ckpt[s], s)):
            extra = full[s] - ckpt[s]
            if total + extra <= hbm_budget:
                keep.add(s)
                total += extra
        groups = []
        a = [s for s in members if s in keep]
        c = [s for s in members if s not in keep]
        if a:
            groups.append((True, 1, a))
        if c:
            groups.append((False, K, c))
        return groups

    if sum(ckpt.values()) <= hbm_budget:
        return [modes(idx)]
    rounds, load = [], []
    for s in sorted(idx, key=lambda s: (-ckpt[s], s)):
        for r, members in enumerate(rounds):
            if load[r] + ckpt[s] <= hbm_budget:
                members.append(s)
                load[r] += ckpt[s]
                break
        else:
            rounds.append([s])
            load.append(ckpt[s])
    return [modes(sorted(m)) for m in rounds]


class PooledGaugeInversion:
    """Independent instances (one mesh each), each with its own checkpoint
    pool (plan_pools): the instances of one tape mode are advanced together
    by one launch per step (a GaugeInversion over a Batch per group), and the
    groups of one admission round have their tapes alive together.

    Gauges are cell indices per instance; eta0 and b are concatenated over
    the instances in instance order, as the series' columns are over their
    gauges.  `forward_backward(eta0, b, loss_fn)` runs round after round and
    returns per-instance results (the loss of each instance only depends on
    its own gauges)."""

    def __init__(self, meshes, gauges, n_steps: int, params, *, checkpoint_every="auto",
                 hbm_budget: float | None = None, device="cuda:0"):
        self.device = torch.device(device)
        self.n_inst = len(meshes)
        self.n_steps = int(n_steps)
        cells = [m.n for m in meshes]
        self.offsets = np.cumsum([0] + cells)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = 0.5 * free
        self.hbm_budget = float(hbm_budget)
        self.rounds = plan_pools(cells, self.n_steps, self.hbm_budget, checkpoint_every)
        g_counts = [len(g) for g in gauges]
        g_offs = np.cumsum([0] + g_counts)
        self.gauge_count = int(g_offs[-1])
        self.groups = []
        for r, groups in enumerate(self.rounds):
            for store_all, K, members in groups:
                cell_idx = np.concatenate([np.arange(self.offsets[s], self.offsets[s + 1])
                                           for s in members])
                local_off = np.cumsum([0] + [cells[s] for s in members])
                g_local = np.concatenate([np.asarray(gauges[s], np.int64) + local_off[i]
                                          for i, s in enumerate(members)])
                cols = np.concatenate([np.arange(g_offs[s], g_offs[s + 1]) for s in members])
                prob = GaugeInversion([meshes[s] for s in members], g_local, n_steps,
                                      [params[s] for s in members],
                                      checkpoint_every="all" if store_all else K,
                                      device=device)
                self.groups.append(dict(
                    round=r, store_all=store_all, K=prob.K, members=list(members), prob=prob,
                    cells=torch.as_tensor(cell_idx, device=self.device),
                    cols=torch.as_tensor(cols, device=self.device),
                    tape_bytes=sum(tape_bytes(cells[s], n_steps, store_all, prob.K)
                                   for s in members)))

    def dt_of(self, instance: int) -> torch.Tensor:
        """The frozen dt sequence of one instance, [n_steps]."""
        for g in self.groups:
            if instance in g["members"]:
                return g["prob"].dt[: self.n_steps, g["members"].index(instance)]
        raise IndexError(instance)

    def plan(self) -> dict:
        """The admission plan: per round, its groups' modes and instances."""
        return {"hbm_budget": self.hbm_budget, "rounds": [
            [{"store_all": g["store_all"], "K": g["K"], "instances": g["members"],
              "tape_bytes": g["tape_bytes"]} for g in self.groups if g["round"] == r]
            for r in range(len(self.rounds))]}

    def record_dt(self, eta0: torch.Tensor, b: torch.Tensor, dt0) -> None:
        dt0 = np.broadcast_to(np.asarray(dt0, np.float64), (self.n_inst,))
        for g in self.groups:
            g["prob"].record_dt(eta0.detach()[g["cells"]].contiguous(),
                                b.detach()[g["cells"]].contiguous(), dt0[g["members"]])

    def gauge_series(self, eta0: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        """The series of every instance (columns in instance and gauge order)
        without keeping any tape (no gradients)."""
        out = torch.empty(self.n_steps, self.gauge_count, dtype=torch.float64, device=self.device)
        with torch.no_grad():
            for g in self.groups:
                out[:, g["cols"]] = g["prob"].gauge_series(eta0[g["cells"]].contiguous(),
                                                           b[g["cells"]].contiguous())
        return out

    def forward_backward(self, make_inputs, loss_fn):
        """make_inputs() -> (eta0, b, leaves): fresh differentiable inputs,
        rebuilt for every round; loss_fn(series, cols, members) -> per-member
        losses.  Runs each round's forward sweeps, then its backward sweep
        (freeing its tapes), and returns ([per-instance loss], grads of
        `leaves` summed over rounds)."""
        losses = torch.zeros(self.n_inst, dtype=torch.float64, device=self.device)
        grads = None
        for r in range(len(self.rounds)):
            eta0, b, leaves = make_inputs()
            total = None
            for g in self.groups:
                if g["round"] != r:
                    continue
                G = g["prob"].gauge_series(eta0[g["cells"]].contiguous(), b[g["cells"]].contiguous())
                per = loss_fn(G, g["cols"], g["members"])
                losses[g["members"]] = per.detach()
                total = per.sum() if total is None else total + per.sum()
            gr = torch.autograd.grad(total, leaves, allow_unused=True)
            gr = [torch.zeros_like(l) if x is None else x for x, l in zip(gr, leaves)]
            grads = gr if grads is None else [a + c for a, c in zip(grads, gr)]
        return losses, grads

    def close(self):
        for g in self.groups:
            g["prob"].close()


class _GaugeSeries(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eta0, b, problem):
        b = b.detach().contiguous()
        series, tape = problem.run_forward(eta0, b)
        ctx.problem = problem
        ctx.tape = tape
        ctx.save_for_backward(b)
        return series

    @staticmethod
    def backward(ctx, g_series):
        (b,) = ctx.saved_tensors
        g_eta0, g_b = ctx.problem.run_backward(b, ctx.tape, g_series.contiguous())
        ctx.tape = None
        return g_eta0, g_b, None
