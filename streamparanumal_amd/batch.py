"""Batched uneven instances: many independent hex meshes, each with K fields,
run as one gather+scatter launch over the block-diagonal union of their
operators and one batched CG update with a per-instance FP64 rdotr, replayed
from a captured graph.  No reference counterpart: the reference runs one mesh
per process and one vector per launch (BS/BS8, BS/BS5); the per-instance
operations are the reference's (ogsKernels.okl:32-81, README.md:49-53).
"""
import numpy as np

from . import _lib
from .device import DeviceArray, Graph, convert, fill_uniform
from .ogs import Ogs
from .streaming import CgPlan, Reducer

_A = 6364136223846793005
_C = 1
_M = (1 << 64) - 1


def lcg_uniform01(count, seed):
    """The reference LCG (libs/primitives/random.cpp:45-48) as U[0,1): draw i -> (x_i >> 11) * 2^-53."""
    x = (_A * (_M ^ seed) + _C) & _M
    out = []
    for _ in range(count):
        out.append((x >> 11) * 2.0**-53)
        x = (_A * x + _C) & _M
    return out


def instances(count=64, seed=4, n_lo=4, n_hi=48, degrees=(3, 5, 7, 9)):
    """Per-side element counts uniform on [n_lo, n_hi] and degrees uniform on `degrees`, from the
    LCG stream of `seed`: instance i uses draws 2i (n) and 2i+1 (p)."""
    u = lcg_uniform01(2 * count, seed)
    span = n_hi - n_lo + 1
    return [(n_lo + min(int(u[2 * i] * span), span - 1), degrees[min(int(u[2 * i + 1] * len(degrees)),
                                                                      len(degrees) - 1)])
            for i in range(count)]


class Batch:
    """Meshes (n, p) of n^3 hex elements at degree p, K fields per node, values stored as `dtype`
    (float64, float32 or uint16 bf16 bits).  parts[i] is instance i's own ogs handle; union is the
    block-diagonal operator over all of them; vectors hold every instance back to back (node
    offsets `offsets`, entry offsets K*offsets)."""

    def __init__(self, meshes, K=3, dtype=np.float32, parts=None, union=None):
        self.meshes = list(meshes)
        self.K = K
        self.dtype = np.dtype(dtype)
        self.storage = _lib.storage_of(self.dtype)
        if parts is None:
            parts = [Ogs.hex(n, n, n, p, keep_ids=False) for (n, p) in self.meshes]
        self.parts = parts
        self.union = union if union is not None else Ogs.batch(self.parts)
        self.offsets = np.zeros(len(self.parts) + 1, dtype=np.int64)
        for i, g in enumerate(self.parts):
            self.offsets[i + 1] = self.offsets[i] + g.N
        self.seg_starts = K * self.offsets
        if np.any(self.seg_starts % 4):
            raise ValueError("every instance needs K*N_local a multiple of 4 (aligned segments)")
        self.N = int(self.offsets[-1])
        self.global_dofs = int(sum(g.NgatherGlobal for g in self.parts))
        self.plan = None
        self.vectors = None

    def allocate(self, seed=5):
        """q (gathered in place, the CG's Ap), p, x, r: U[-1,1) from seeds seed..seed+3."""
        n = self.K * self.N
        vs = []
        tmp = DeviceArray(n)
        for s in range(4):
            fill_uniform(tmp, seed + s, 0)
            if self.dtype == np.float64:
                v = DeviceArray(n)
                from .streaming import copy
                copy(tmp, v)
            else:
                v = DeviceArray(n, self.dtype)
                convert(tmp, v)
            vs.append(v)
        del tmp
        self.vectors = dict(zip(("q", "p", "x", "r"), vs))
        return self.vectors

    def autotune(self, reps=5):
        """Pick the gather+scatter launch shape and the CG plan by timing them on this batch."""
        v = self.vectors
        gs = self.union.autotune(self.dtype, self.K, reps, mixed=True)
        self.plan = CgPlan.autotuned(self.seg_starts, v["p"], v["q"], v["x"], v["r"], reps)
        return {"gather_scatter": gs, "cg": self.plan.tuning}

    def sibling(self, dtype):
        """The same instances and union operator with values stored as another dtype."""
        return Batch(self.meshes, self.K, dtype, parts=self.parts, union=self.union)

    def make_plan(self, target_blocks=2**31 - 1):
        self.plan = CgPlan(self.seg_starts, target_blocks)

    def step(self, alpha):
        """q <- GS(q) for every instance (one launch), then x += alpha p, r -= alpha q, rdotr per instance."""
        v = self.vectors
        self.union.gather_scatter_mixed(v["q"], self.K)
        self.plan.update(alpha, v["p"], v["q"], v["x"], v["r"])

    def capture(self, alpha):
        """The step as one replayable graph."""
        return Graph.capture(lambda: self.step(alpha))


class SeparateFP64:
    """The plain form: one FP64 gather+scatter and one FP64 CG update launch sequence per instance,
    each on its own vectors, through the instance's own ogs handle."""

    def __init__(self, batch, seed=5):
        self.b = batch
        self.vecs = []
        self.red = []
        for i, g in enumerate(batch.parts):
            n = batch.K * g.N
            vs = []
            for s in range(4):
                v = DeviceArray(n)
                fill_uniform(v, seed + s, int(batch.seg_starts[i]))  # the same draws as the batch holds
                vs.append(v)
            self.vecs.append(dict(zip(("q", "p", "x", "r"), vs)))
            self.red.append(Reducer(n, _lib.REDUCE_REFERENCE, 256))

    def step(self, alpha):
        for g, v, red in zip(self.b.parts, self.vecs, self.red):
            g.gather_scatter(v["q"], self.b.K)
            red.cg_full(alpha, v["p"], v["q"], v["x"], v["r"])

    def rdotr(self):
        return np.array([r.value() for r in self.red])
