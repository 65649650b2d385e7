"""Gradient of a dye observation misfit with respect to the source loads and
the Manning n field, by the discrete adjoint of the hot-path step.

The forward model is the device-resident step of ``Context.step`` (closure,
CALPUV2C, closure, CALCONC) on one basin; the misfit is

    J = sum over observation steps t and observed cells o of
        ( sum_K DZC(o,K) DYE(o,K,t) - d(o,t) )^2

(the layer-mean dye at cell o after step t).  ``loads[t, s]`` is the load of
source record s during step t+1: its CQS on every layer is
``loads[t, s] / Q(s)`` with Q(s) = sum_K QSS(s,K), so the mass rate per
layer is loads * QSS(s,K) / Q(s).

``gradient`` runs the forward model once, saving the whole context state
every ``ckpt`` steps (efdc_state_save), then walks the segments backwards:
each segment is re-run from its checkpoint, keeping after every step the
state and the step's recording (what the reverse reads: the CG solution,
its scaling, the old constituent levels; efdc_adj_record_save) in one
buffer, and every step is reversed by loading both and calling
efdc_adj_step.  Two forward evaluations per step in all.  Memory:
(nsteps/ckpt + ckpt) snapshots, device buffers from efdc_state_alloc kept
in a pool for the life of the problem.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from . import lib


def coarse_index(grid, nxc: int, nyc: int) -> np.ndarray:
    """For every cell L (index L-1), the index of the block of an
    nxc x nyc coarsening of the IC-2 x JC-2 interior that holds it (-1 on
    ghosts)."""
    idx = np.full(grid.LCM, -1, dtype=np.int64)
    L = np.arange(2, grid.LA + 1) - 1
    ni, nj = grid.IC - 2, grid.JC - 2
    bi = ((grid.IL[L] - 2) * nxc) // ni
    bj = ((grid.JL[L] - 2) * nyc) // nj
    idx[L] = bj * nxc + bi
    return idx


class DyeMisfit:
    """Misfit J(loads, n) of one basin and its gradient."""

    def __init__(self, case, *, nsteps: int, obs_L, obs_every: int, data=None,
                 dye_slot: int = 1, ckpt: int = 20, store: int = lib.STORE_FP64,
                 cg_rel_tol: float = 1e-12, **params):
        self.case = case
        self.nsteps = nsteps
        self.obs_L = np.asarray(obs_L, np.int32)
        self.obs_every = obs_every
        self.dye = dye_slot
        self.ckpt = max(1, ckpt)
        g = case.grid
        self.KC = g.KC
        params.setdefault("RSQM", 0.0)
        params.setdefault("cg_rel_tol", 1e-13)
        params.setdefault("ITERM", 4000)
        self.ctx = lib.Context(g, store=store, **params)
        self.ctx.upload_case(case)
        self.ctx.adj_enable([dye_slot], cg_rel_tol=cg_rel_tol)
        src = case.sources
        self.nsrc = len(src["LQS"])
        self.qtot = np.asarray(src["qss"]).reshape(self.nsrc, self.KC).sum(axis=1)
        self.cqs0 = np.array(src["cqs"], np.float64).reshape(self.nsrc, self.KC, -1)
        self.nc = self.cqs0.shape[2]
        self.dzc_obs = case.k["DZC"][:self.KC, self.obs_L - 1]      # (KC, nobs)
        self._obs_K = np.repeat(np.arange(1, self.KC + 1), len(self.obs_L)).astype(np.int32)
        self._obs_LL = np.tile(self.obs_L, self.KC).astype(np.int32)
        self.obs_steps = [t for t in range(1, nsteps + 1) if t % obs_every == 0]
        self.data = None if data is None else np.asarray(data, np.float64)
        self.nman0 = None if case.manning_n is None else np.array(case.manning_n, np.float64)
        self._pool = []          # snapshot buffers (device pointers)
        self._state_bytes = self.ctx.state_bytes()
        self._init = self.ctx.state_alloc()
        self.ctx.state_save(self._init)
        self.ctx.synchronize()

    # ------------------------------------------------------------ helpers
    def _buffers(self, n):
        """n snapshot buffers from the pool (grown on demand)."""
        while len(self._pool) < n:
            self._pool.append(self.ctx.state_alloc())
        return self._pool[:n]

    def _cqs(self, loads, n):
        c = self.cqs0.copy()
        c[:, :, self.dye] = (loads[n - 1] / self.qtot)[:, None]
        return c

    def _observe(self):
        c = self.ctx.get_constituent_at(self.dye, self._obs_LL, self._obs_K)
        return np.sum(self.dzc_obs * c.reshape(self.KC, len(self.obs_L)), axis=0)

    def _step(self, loads, n):
        self.ctx.set_source_cqs(self._cqs(loads, n))
        self.ctx.step(self.case.dt, n)

    def _start(self, loads, nman):
        self.ctx.state_load(self._init)
        if nman is None:
            nman = self.nman0
        if nman is not None:
            self.ctx.set_manning(nman)

    # ------------------------------------------------------------ forward
    def observations(self, loads, nman=None) -> np.ndarray:
        """Observed layer-mean dye (len(obs_steps), nobs) of one forward run."""
        loads = np.asarray(loads, np.float64)
        self._start(loads, nman)
        out = []
        for n in range(1, self.nsteps + 1):
            self._step(loads, n)
            if n % self.obs_every == 0:
                out.append(self._observe())
        return np.array(out)

    def misfit(self, loads, nman=None) -> float:
        y = self.observations(loads, nman)
        return float(np.sum((y - self.data) ** 2))

    # ------------------------------------------------------------ reverse
    def gradient(self, loads, nman=None, *, rerun_steps=False):
        """(J, dJ/dloads (nsteps, nsrc), dJ/dn per cell (LCM,), stats).

        rerun_steps: instead of keeping each step's recording, reverse
        every step by loading the state before it and running it again
        (three forward evaluations per step; the same bits)."""
        loads = np.asarray(loads, np.float64)
        ctx, K = self.ctx, self.ckpt
        rec_off = self._state_bytes
        ctx.adj_reset()
        self._start(loads, nman)
        starts = list(range(0, self.nsteps, K))
        bufs = self._buffers(len(starts) + min(K, self.nsteps))
        ckpts = dict(zip(starts, bufs[:len(starts)]))
        seg = bufs[len(starts):]
        resid = {}
        for n in range(1, self.nsteps + 1):
            if (n - 1) % K == 0:
                ctx.state_save(ckpts[n - 1])
            self._step(loads, n)
            if n % self.obs_every == 0:
                resid[n] = self._observe() - self.data[n // self.obs_every - 1]
        J = float(sum(np.sum(r ** 2) for r in resid.values()))
        gl = np.zeros((self.nsteps, self.nsrc))
        iters = []
        for t0 in sorted(ckpts, reverse=True):
            t1 = min(t0 + K, self.nsteps)
            ctx.state_load(ckpts[t0])
            for j, n in enumerate(range(t0 + 1, t1 + 1)):
                if rerun_steps:
                    # the state before step n
                    ctx.state_save(seg[j])
                    if n < t1:
                        self._step(loads, n)
                else:
                    # the state after step n and what its reverse reads
                    self._step(loads, n)
                    ctx.state_save(seg[j])
                    ctx.adj_record_save(seg[j] + rec_off)
            for n in range(t1, t0, -1):
                ctx.state_load(seg[n - t0 - 1])
                if rerun_steps:
                    self._step(loads, n)
                else:
                    ctx.adj_record_load(seg[n - t0 - 1] + rec_off)
                if n in resid:
                    r = resid[n]
                    kk = np.repeat(np.arange(1, self.KC + 1), len(self.obs_L))
                    LL = np.tile(self.obs_L, self.KC)
                    val = (2.0 * self.dzc_obs * r[None, :]).reshape(-1)
                    ctx.adj_seed_con(self.dye, LL, kk, val)
                st = ctx.adj_step(self.case.dt, n)
                iters.append(st["cg_iters"])
                g = ctx.adj_source_grad(self.nsrc)[:, :, self.dye]
                gl[n - 1] = g.sum(axis=1) / self.qtot
        gn = ctx.adj_get(lib.ADJ_MANNING)
        ctx.synchronize()
        return J, gl, gn, dict(adj_cg_iters=iters)

    def close(self):
        if self.ctx.h:
            for b in self._pool + [self._init]:
                self.ctx.state_free(b)
            self._pool = []
        self.ctx.close()


def autograd_function():
    """A torch.autograd.Function over (loads, coarse n) -> J for a DyeMisfit
    problem: forward runs the model, backward runs the adjoint."""
    import torch

    class Misfit(torch.autograd.Function):
        @staticmethod
        def forward(fctx, loads, ncoarse, prob, cidx):
            nman = ncoarse.detach().cpu().numpy().reshape(-1)[np.maximum(cidx, 0)]
            fctx.prob, fctx.cidx = prob, cidx
            fctx.save_for_backward(loads, ncoarse)
            J = prob.misfit(loads.detach().cpu().numpy(), nman)
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(fctx, gJ):
            loads, ncoarse = fctx.saved_tensors
            prob, cidx = fctx.prob, fctx.cidx
            nman = ncoarse.detach().cpu().numpy().reshape(-1)[np.maximum(cidx, 0)]
            _, gl, gn, _ = prob.gradient(loads.detach().cpu().numpy(), nman)
            gc = np.zeros(ncoarse.numel())
            m = cidx >= 0
            np.add.at(gc, cidx[m], gn[m])
            return (gJ * torch.from_numpy(gl), gJ * torch.from_numpy(gc).reshape(ncoarse.shape),
                    None, None)

    return Misfit
