"""Batched imaginary-time ground-state sweep: many independent instances of
uneven grid sizes, one group (batched launches) per grid shape, the groups'
heavy-ball iterations captured as hipGraphs (one per group, each on a stream
of its own, or one graph of all groups on a shared stream) and replayed until
every instance has met |E - E_prev| / E_prev < eps (the reference's test,
src/oneComponentGPSolver.cu:84-88) or max_iters is reached."""
from __future__ import annotations

import time

import numpy as np

from ._lib import FGPE_PREC_MIXED, Context, Graph, Group


class Sweep:
    def __init__(self, instances, precision=FGPE_PREC_MIXED, stream=None, new_stream=None):
        """stream: a HIP stream handle shared by every context (a graph cannot
        capture the null stream); all groups then replay one graph together.
        new_stream: instead, a callable returning a fresh stream handle, called
        once per grid shape: each group runs on a stream of its own and run()
        replays one graph per group, each at its own pace (the caller keeps
        the streams alive)."""
        self.instances = list(instances)
        self.ctxs = []
        shapes = {}
        self.streams = None if new_stream is None else {}
        for inst in self.instances:
            g = inst.grid
            c = Context(g.nx, g.ny, g.nz, precision)
            key = (g.nx, g.ny, g.nz)
            if new_stream is not None:
                if key not in self.streams:
                    self.streams[key] = new_stream()
                c.set_stream(self.streams[key])   # before any work is queued on the context
            elif stream is not None:
                c.set_stream(stream)
            kx, ky, kz = g.kaxes()
            c.set_kspace(kx, ky, kz, g.dv)
            c.set_hamiltonian(a=(inst.a,), a_dd=inst.a_dd, dip=inst.dip, sc_U=inst.sc_U, sc_R=inst.sc_R,
                              sc_dim=inst.sc_dim)
            c.upload_psi(inst.initial_state())
            c.upload_vext(inst.trap())
            self.ctxs.append(c)
            shapes.setdefault(key, []).append(len(self.ctxs) - 1)
        self.members = list(shapes.values())        # instance indices per group
        self.shapes = list(shapes.keys())
        self.groups = [Group([self.ctxs[i] for i in m]) for m in self.members]
        self.graph = None

    def autotune(self, reps=3):
        return {s: g.autotune(reps) for s, g in zip(self.shapes, self.groups)}

    def run(self, dt, beta, eps, iters_per_graph=50, max_iters=20000, sync=None, recapture=0.75):
        """Replays captured iterations until all instances converge (or
        max_iters iterations have been replayed).  Converged instances only
        cost an early exit per block, but their blocks are still launched:
        once a `recapture` share of the instances in the current graph has
        converged, the graph is captured again over the unconverged ones only
        (each heavy-ball run continues where it stopped).  Returns
        per-instance iteration counts, convergence flags, the wall time of the
        replays (captures included) and the number of replays."""
        if self.streams is not None:
            return self._run_per_group(dt, beta, eps, iters_per_graph, max_iters, sync, recapture)
        n = len(self.instances)
        iters = np.zeros(n, dtype=np.int64)
        conv = np.zeros(n, dtype=bool)
        launches = replayed = 0
        self.captures = 0
        active = [list(m) for m in self.members]
        t0 = time.perf_counter()
        done = False
        while not done and replayed < max_iters:
            members = [m for m in active if m]
            groups = [Group([self.ctxs[i] for i in m]) for m in members]
            graph = Graph(groups, dt, iters_per_graph, beta=beta, npart=1.0, eps=eps)
            self.captures += 1
            base = iters.copy()   # the capture restarts every context's iteration counter
            n_active = sum(len(m) for m in members)
            while replayed < max_iters:
                graph.launch()
                launches += 1
                replayed += iters_per_graph
                n_conv = 0
                for m, g in zip(members, groups):
                    it, cv = g.imag_status()
                    iters[m] = base[m] + it
                    conv[m] = cv
                    n_conv += int(cv.sum())
                if n_conv == n_active:
                    done = True
                    break
                if n_conv >= max(1.0, recapture * n_active):
                    break
            graph.close()
            for g in groups:
                g.close()
            active = [[i for i in m if not conv[i]] for m in active]
        if sync is not None:
            sync()
        wall = time.perf_counter() - t0
        for g in self.groups:
            g.imag_finish()   # applies every instance's pending renormalisation
        return iters, conv, wall, launches

    def _run_per_group(self, dt, beta, eps, iters_per_graph, max_iters, sync, recapture):
        """run() with one graph per group, each group on its own stream: the
        host waits for one group's replay at a time and relaunches it at once,
        while the other groups' replays go on; a group recaptures over its own
        unconverged instances and stops when they have all converged (or after
        max_iters iterations of its own)."""
        n = len(self.instances)
        iters = np.zeros(n, dtype=np.int64)
        conv = np.zeros(n, dtype=bool)
        self.captures = 0
        launches = 0
        st = [{"active": list(m), "done": False, "replayed": 0} for m in self.members]

        def capture(s):
            s["group"] = Group([self.ctxs[i] for i in s["active"]])
            s["graph"] = Graph([s["group"]], dt, iters_per_graph, beta=beta, npart=1.0, eps=eps)
            s["base"] = iters.copy()   # the capture restarts every context's iteration counter
            self.captures += 1

        def release(s):
            s["graph"].close()
            s["group"].close()

        t0 = time.perf_counter()
        for s in st:
            capture(s)
        for s in st:
            s["graph"].launch()
            s["replayed"] += iters_per_graph
            launches += 1
        while not all(s["done"] for s in st):
            for s in st:
                if s["done"]:
                    continue
                m = s["active"]
                it, cv = s["group"].imag_status()   # waits for this group's stream only
                iters[m] = s["base"][m] + it
                conv[m] = cv
                n_conv = int(cv.sum())
                if n_conv == len(m) or s["replayed"] >= max_iters:
                    release(s)
                    s["done"] = True
                    s["finish_s"] = time.perf_counter() - t0
                    continue
                if n_conv >= max(1.0, recapture * len(m)):
                    release(s)
                    s["active"] = [i for i in m if not conv[i]]
                    capture(s)
                s["graph"].launch()
                s["replayed"] += iters_per_graph
                launches += 1
        if sync is not None:
            sync()
        wall = time.perf_counter() - t0
        self.group_finish_s = [s["finish_s"] for s in st]   # when each group's last replay was seen done
        for g in self.groups:
            g.imag_finish()
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
