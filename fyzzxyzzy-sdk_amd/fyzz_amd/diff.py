"""The differentiable form: a rollout of independent scenes whose final
positions are differentiated with respect to the initial linear and angular
velocities and each scene's friction coefficient.

Forward: the device world steps the scenes (per-scene broadphase, PCM, Dy)
and records each step's contact set (the Dy managers and the friction
patches they read) plus the body records every `checkpoint_every` steps.
Backward: for each segment between checkpoints, newest first, the Dy steps
are recomputed from the checkpoint with the recorded contact sets (the Dy
step is deterministic, so they reproduce the forward bit for bit), then
the steps are reversed newest first: each is recomputed once more with the
tape on and its adjoint runs (fyzz_hip_dy_adjoint: PGS and integration with
the step's contact rows frozen).
"""
from __future__ import annotations

import numpy as np

from .dy import DY_BODY, DY_STATIC, DySolver, to_device
from .world import World


class Rollout:
    def __init__(self, ctx, half, pos, quat, scene_counts, steps, dt=1.0 / 120.0, checkpoint_every=16,
                 floor_friction=0.5, density=1000.0):
        self.ctx = ctx
        self.half = np.ascontiguousarray(half, np.float32)
        self.pos = np.ascontiguousarray(pos, np.float32)
        self.quat = np.ascontiguousarray(quat, np.float32)
        self.counts = np.ascontiguousarray(scene_counts, np.uint32)
        self.nb = int(self.counts.sum())
        self.steps = int(steps)
        self.dt = float(dt)
        self.every = int(checkpoint_every)
        self.floor_friction = float(floor_friction)
        self.density = float(density)
        self.scene_of_body = np.repeat(np.arange(len(self.counts)), self.counts)
        st = np.zeros(1, DY_STATIC)
        st["orientation"][0, 3] = 1.0
        st["static_friction"] = st["dynamic_friction"] = floor_friction
        self.statics = to_device(st)
        self.solver = None
        self.records = []
        self.checkpoints = {}

    def forward(self, v0, w0, mu_body):
        """v0, w0: (nb, 3) host or device arrays; mu_body: (nb,) friction.
        Returns the final positions (nb, 3) as a device float32 tensor.
        Each step's contact set is appended on the device to arenas sized
        for 4 managers per body and step; the counts are read once, after
        the rollout (a rollout that needs more runs again with exact
        arenas)."""
        return self._forward(v0, w0, mu_body, None)

    def _forward(self, v0, w0, mu_body, capacity):
        import torch
        from .dy import DY_MANAGER, DY_PATCH
        v0_in, w0_in, mu_in = v0, w0, mu_body
        v0 = np.ascontiguousarray(torch.as_tensor(v0).detach().cpu().numpy(), np.float32)
        w0 = np.ascontiguousarray(torch.as_tensor(w0).detach().cpu().numpy(), np.float32)
        mu = np.ascontiguousarray(torch.as_tensor(mu_body).detach().cpu().numpy(), np.float32)
        if capacity is None:
            capacity = self.steps * max(4 * self.nb, 1024)
        mg_arena = torch.empty((max(capacity, 1), DY_MANAGER.itemsize), dtype=torch.uint8, device="cuda")
        hs_arena = torch.empty((max(capacity, 1), DY_PATCH.itemsize), dtype=torch.uint8, device="cuda")
        cursor = torch.zeros(1, dtype=torch.int64, device="cuda")
        counts = torch.zeros(self.steps, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()  # the zeroed cursor and counts before the world's stream uses them
        w = World(self.ctx, self.nb + 1)
        fid = w.add_static_floor(0.0)
        w.set_body_material(fid, self.floor_friction, self.floor_friction, 0.0)
        first = w.add_dynamic_boxes(self.half, self.pos, self.quat, linear=v0, angular=w0, density=self.density)
        w.set_materials(first, mu, mu, np.zeros(self.nb, np.float32))
        w.set_scenes(self.counts)
        self.records = []
        self.checkpoints = {}
        for s in range(self.steps):
            if s % self.every == 0:
                self.checkpoints[s] = w.copy_bodies().clone()
            w.step(self.dt)
            w.append_dy_inputs(mg_arena, hs_arena, capacity, cursor, counts[s:s + 1])
        final = w.copy_bodies().clone()
        st = w.stats()  # synchronizes the world's stream
        w.close()
        if st["flags"]:
            raise RuntimeError(f"rollout status flags {st['flags']:#x}")
        n = counts.cpu().numpy().astype(np.int64)
        if int(cursor.item()) == -1:  # the arenas were too small: once more with the exact size
            return self._forward(v0_in, w0_in, mu_in, int(n.sum()))
        off = np.concatenate([[0], np.cumsum(n)])
        self.records = [(mg_arena[off[k]:off[k + 1]], hs_arena[off[k]:off[k + 1]]) for k in range(self.steps)]
        return final[:, :12].contiguous().view(torch.float32).reshape(self.nb, 3)

    def close(self):
        """Releases the adjoint solver's device workspace."""
        if self.solver is not None:
            self.solver.close()
            self.solver = None

    def _dy_step(self, bodies, s, tape):
        self.solver.set_tape(tape)
        mg, hs = self.records[s]
        self.solver.step(self.dt, bodies, self.statics, mg, hs if len(hs) else None)

    def backward(self, adj_pos):
        """Adjoint of the final positions in; (grad v0, grad w0, grad of
        each body's friction) out, device tensors."""
        import torch
        if self.solver is None:
            cap = max(max((len(m) for m, _ in self.records), default=1), 1)
            self.solver = DySolver(self.ctx, self.nb, cap)
        elif max((len(m) for m, _ in self.records), default=1) > self.solver_cap:
            self.solver.close()
            self.solver = DySolver(self.ctx, self.nb, max(len(m) for m, _ in self.records))
        self.solver_cap = max((len(m) for m, _ in self.records), default=1)
        dev = adj_pos.device
        ap = adj_pos.detach().to(torch.float32).contiguous()
        av = torch.zeros((self.nb, 3), dtype=torch.float32, device=dev)
        aw = torch.zeros_like(av)
        amu = torch.zeros(self.nb, dtype=torch.float32, device=dev)
        starts = sorted(self.checkpoints)
        for seg_start in reversed(starts):
            seg_end = min(seg_start + self.every, self.steps)
            states = [self.checkpoints[seg_start].clone()]
            for s in range(seg_start, seg_end - 1):
                b = states[-1].clone()
                self._dy_step(b, s, tape=False)
                states.append(b)
            for s in reversed(range(seg_start, seg_end)):
                b = states[s - seg_start].clone()
                self._dy_step(b, s, tape=True)
                av, aw, af = self.solver.adjoint(av, aw, ap)
                amu += af
            self.solver.status()
        return av, aw, amu


def rollout_function():
    """torch.autograd.Function over Rollout: (v0, w0, mu per scene) ->
    final positions."""
    import torch

    class RolloutFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, w0, mu_scene, rollout):
            scene = torch.as_tensor(rollout.scene_of_body, device=mu_scene.device)
            mu_body = mu_scene.detach()[scene]
            out = rollout.forward(v0, w0, mu_body)
            ctx.rollout = rollout
            # scenes own contiguous body ranges: per-scene sums are segment
            # sums (a fixed order, unlike atomic index adds)
            ctx.lengths = torch.bincount(scene, minlength=mu_scene.shape[0])
            return out

        @staticmethod
        def backward(ctx, grad_pos):
            gv, gw, gmu_body = ctx.rollout.backward(grad_pos)
            gmu = torch.segment_reduce(gmu_body, "sum", lengths=ctx.lengths)
            return gv, gw, gmu, None

    return RolloutFn
