"""Batched imaginary-time ground-state sweep: many independent instances of
uneven grid sizes, one group (batched launches) per grid shape, all groups'
heavy-ball iterations captured in one hipGraph and replayed until every
instance has met |E - E_prev| / E_prev < eps (the reference's test,
src/oneComponentGPSolver.cu:84-88) or max_iters is reached."""
from __future__ import annotations

import time

import numpy as np

from ._lib import FGPE_PREC_MIXED, Context, Graph, Group


class Sweep:
    def __init__(self, instances, precision=FGPE_PREC_MIXED, stream=None):
        """stream: a HIP stream handle shared by every context (a graph cannot
        capture the null stream)."""
        self.instances = list(instances)
        self.ctxs = []
        shapes = {}
        for inst in self.instances:
            g = inst.grid
            c = Context(g.nx, g.ny, g.nz, precision)
            if stream is not None:
                c.set_stream(stream)
            kx, ky, kz = g.kaxes()
            c.set_kspace(kx, ky, kz, g.dv)
            c.set_hamiltonian(a=(inst.a,), a_dd=inst.a_dd, dip=inst.dip, sc_U=inst.sc_U, sc_R=inst.sc_R,
                              sc_dim=inst.sc_dim)
            c.upload_psi(inst.initial_state())
            c.upload_vext(inst.trap())
            self.ctxs.append(c)
            shapes.setdefault((g.nx, g.ny, g.nz), []).append(len(self.ctxs) - 1)
        self.members = list(shapes.values())        # instance indices per group
        self.shapes = list(shapes.keys())
        self.groups = [Group([self.ctxs[i] for i in m]) for m in self.members]
        self.graph = None

    def autotune(self, reps=3):
        return {s: g.autotune(reps) for s, g in zip(self.shapes, self.groups)}

    def run(self, dt, beta, eps, iters_per_graph=50, max_iters=20000, sync=None):
        """Replays the captured iterations until all instances converge.
        Returns per-instance iteration counts, convergence flags, the wall
        time of the replays and the number of replays."""
        self.graph = Graph(self.groups, dt, iters_per_graph, beta=beta, npart=1.0, eps=eps)
        n = len(self.instances)
        iters = np.zeros(n, dtype=np.int64)
        conv = np.zeros(n, dtype=bool)
        launches = 0
        t0 = time.perf_counter()
        while launches * iters_per_graph < max_iters:
            self.graph.launch()
            launches += 1
            conv[:] = True
            for m, g in zip(self.members, self.groups):
                it, cv = g.imag_status()
                iters[m] = it
                conv[m] = cv
            if conv.all():
                break
        if sync is not None:
            sync()
        wall = time.perf_counter() - t0
        for m, g in zip(self.members, self.groups):
            it, cv = g.imag_finish()
            iters[m] = it
            conv[m] = cv
        return iters, conv, wall, launches

    def energies(self):
        """(norm, mu, E) of every instance's current state."""
        return np.array([c.observables(0.0) for c in self.ctxs])

    def close(self):
        if self.graph is not None:
            self.graph.close()
        for g in self.groups:
            g.close()
        for c in self.ctxs:
            c.close()
