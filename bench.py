"""Benchmark: real-time RK4 steps/s of the dipolar GPE on one MI355X.

Workload (configurable with --grid): 256^3 one-component dipolar GPE, box
16 x 16 x 32 (oscillator units), cigar trap omega = (1, 1, 0.25), g = 200,
eps_dd = 1.3 with dipoles along z, dt = 5e-4; psi0 = seeded Thomas-Fermi
profile + 1e-3 complex noise.  Synthetic data generated in-process.

`value` is the plain FP64 form's steps/s with psi and V resident in HBM; the
mixed-precision form is timed the same way and reported beside it with its
error against the FP64 form.  The roofline object covers the kernel class
with the largest share of the FP64 step time, from HIP events recorded on
the stream the kernels run on, over a second timed region of the same length.

The same line carries the metric's second part, adjoint gradients/s
(`adjoint`: 128^3 dipolar, 2000 RK4 steps, checkpoint every 45, both forms,
the FP64 gradient checked by central differences along 4 seeded V
directions, a and a_dd), and the oracle's steps/s on the host
(`cpu_baseline`: many threads after a warm-up, and one thread).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


PMC_SUMMARY = os.path.join("profiles", "rk4_dipolar_256cubed_pmc_traffic.md")


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    from the committed rocprofv3 summary of this workload, or None."""
    path = os.path.join(ROOT, PMC_SUMMARY)
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch (.+) = (\d+)$", line.strip())
        if m and m.group(1) == kernel:
            return int(m.group(2))
    return None


def run_form(torch, w, psi, V, precision, steps, warmup):
    from fgpe_amd import Context
    g = w.grid
    kx, ky, kz = g.kaxes()
    ctx = Context(g.nx, g.ny, g.nz, precision)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    ctx.set_kspace(kx, ky, kz, g.dv)
    ctx.set_hamiltonian(a=w.a, a_dd=w.a_dd, dip=w.dip)
    ctx.upload_psi(psi)
    ctx.upload_vext(V)
    tune = ctx.autotune(5)
    t = 0.0
    ctx.rk4_steps(w.dt, warmup, t)
    t += warmup * w.dt
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    ctx.rk4_steps(w.dt, steps, t)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    t += steps * w.dt
    # second timed region: per-kernel-class HIP events on the same stream
    ctx.timing_enable(True)
    ctx.rk4_steps(w.dt, steps, t)
    kt = ctx.timing_read()
    ctx.timing_enable(False)
    t += steps * w.dt
    obs = ctx.observables(t)
    out = ctx.download_psi()
    la = ctx.kmul_axis()
    axis_of = {"xstage": 0, "strided": 3 - la if g.nz > 1 else 1, "last": la}
    chosen = {k: v for k, v in ctx.variants().items() if k[1] == axis_of[k[0]]}
    nslab = ctx.slabs()
    order = ctx.order()
    ctx.close()
    return ms / steps, kt, obs, out, {"autotune_ms": {k: [round(x, 4) for x in v] for k, v in tune.items()},
                                      "variant": {k[0]: v for k, v in chosen.items()}, "slabs": nslab,
                                      "order": order, "kmul_axis": la}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, psi, V, steps, warmup, threads, steps_1t):
    """The oracle (oracle/fgpe_oracle.c, its own FFT: no FFTW on these
    machines) on the host: `warmup` untimed steps, then `steps` timed steps
    on `threads` OpenMP threads, then `steps_1t` timed steps on one thread,
    each continuing from the state the previous ones reached."""
    from oracle import oracle as orc
    g = w.grid
    kx, ky, kz = g.kaxes()
    og = orc.Grid(g.nx, g.ny, g.nz, kx, ky, kz, g.dv)
    h = orc.Ham(1, w.a, vtilde=orc.vdd(og, w.a_dd, w.dip))
    orc.set_threads(threads)
    used = orc.num_threads()
    t_start = time.perf_counter()
    t = 0.0
    if warmup:
        psi = orc.rk4(og, h, psi, V, w.dt, warmup, t0=t)
        t += warmup * w.dt
    t0 = time.perf_counter()
    psi = orc.rk4(og, h, psi, V, w.dt, steps, t0=t)
    many = steps / (time.perf_counter() - t0)
    t += steps * w.dt
    one = None
    if steps_1t > 0:
        orc.set_threads(1)
        t0 = time.perf_counter()
        psi = orc.rk4(og, h, psi, V, w.dt, steps_1t, t0=t)
        one = steps_1t / (time.perf_counter() - t0)
        orc.set_threads(threads)
    return many, used, one, time.perf_counter() - t_start


def adjoint_leg(torch, nsteps, every, ndirs):
    """Gradient of the final-state fidelity |<target|psi(T)>|^2 with respect to
    the trap V, a and a_dd through the discrete adjoint, for the dipolar GPE
    (omega = (1,1,0.25), g = 200, eps_dd = 1.3, box 16 x 16 x 32) on 128^3 over
    nsteps RK4 steps of dt = 5e-4 with a checkpoint every `every` steps.  One
    gradient = the checkpointed forward run plus the backward sweep; each
    form is timed twice and the second (warm) time is kept.  The FP64
    gradient is checked by central differences of the FP64 forward along
    `ndirs` of the seeded smooth V directions, a and a_dd; the mixed gradient
    against the FP64 one (finite differences of an FP32-bulk forward are
    dominated by its rounding at these step sizes)."""
    from fgpe_amd import FGPE_PREC_F64, FGPE_PREC_MIXED, Context, configs, roofline
    torch.cuda.set_device(0)
    n, dt = nsteps, 5e-4
    grid = configs.GridSpec(128, 128, 128, 16.0, 16.0, 32.0)
    g = 200.0
    a, a_dd = g / (4 * configs.PI), 1.3 * g / (4 * configs.PI)
    omega = (1.0, 1.0, 0.25)
    V = configs.harmonic_trap(grid, omega).real
    psi0 = configs.thomas_fermi_noisy(grid, 4, g, omega)
    target = configs.thomas_fermi_noisy(grid, 5, g, (1.0, 1.0, 0.3))
    kx, ky, kz = grid.kaxes()
    dv = grid.dv

    import functools

    @functools.lru_cache(maxsize=None)
    def smooth(seed):
        u = configs.splitmix64_uniform(seed, grid.npoints).reshape(grid.shape)
        F = np.fft.fftn(u)
        KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
        f = np.fft.ifftn(F * np.exp(-4.0 * (KX ** 2 + KY ** 2 + KZ ** 2))).real
        return f / np.abs(f).max()

    def make(prec, Vx, ax, addx):
        c = Context(128, 128, 128, prec)
        c.set_kspace(kx, ky, kz, dv)
        c.set_hamiltonian(a=(ax,), a_dd=addx, dip=(0.0, 0.0, 1.0))
        c.upload_vext(Vx.astype(np.complex128))
        c.upload_psi(psi0)
        c.autotune(3)   # plan variants and pass order for this grid (the state is left as uploaded)
        return c

    def J(psiT):
        ov = np.sum(np.conj(target) * psiT) * dv
        return float(abs(ov) ** 2), ov

    out = {}
    for name, prec in (("f64", FGPE_PREC_F64), ("mixed", FGPE_PREC_MIXED)):
        times = []
        for rep in range(2):
            c = make(prec, V, a, a_dd)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            c.adjoint_forward(dt, n, every)
            psiT = c.download_psi()
            Jv, ov = J(psiT)
            lamT = torch.from_numpy(2 * ov * dv * target).cuda()
            gV = torch.empty(grid.shape, dtype=torch.float64, device="cuda")
            lam0 = torch.empty_like(lamT)
            torch.cuda.synchronize()
            gs = c.adjoint_backward(lamT.data_ptr(), lam0.data_ptr(), gV.data_ptr())
            c.sync()
            times.append(time.perf_counter() - t0)
            c.close()
        gVn = gV.cpu().numpy()
        bpp = roofline.adjoint_bytes_per_point(prec == FGPE_PREC_MIXED, n, every, 128)
        gbs = bpp * grid.npoints / times[-1] / 1e9
        res = {"grads_per_s": round(1.0 / times[-1], 4), "s_per_gradient": round(times[-1], 3), "J": Jv,
               "bytes_per_point": round(bpp, 1), "GBs": round(gbs, 1),
               "frac_of_peak": round(gbs / roofline.HBM_PEAK_GBS, 4)}
        print(f"# adjoint {name}: J = {Jv:.10f}, {times[-1]:.2f} s per gradient, {gbs:.0f} GB/s", file=sys.stderr,
              flush=True)
        if name == "mixed":
            ref = out["f64"]
            cmp = [("V", float(np.sum(gVn * smooth(4 + j))), ref["directional"][j]) for j in range(ndirs)]
            cmp += [("a", gs[0], ref["directional"][ndirs]), ("a_dd", gs[1], ref["directional"][ndirs + 1])]
            checks = [{"dir": lab, "mixed": m, "f64": r, "rel": abs(m - r) / abs(r)} for lab, m, r in cmp]
            for x in checks:
                print(f"#   {x['dir']}: mixed {x['mixed']:.10e} f64 {x['f64']:.10e} rel {x['rel']:.2e}",
                      file=sys.stderr, flush=True)
            res.update(vs_f64_adjoint=checks, max_rel_vs_f64=max(x["rel"] for x in checks))
            out[name] = res
            continue
        checks = []
        h = 1e-4
        dirs = [("V", smooth(4 + j), 0.0, 0.0) for j in range(ndirs)] + [
            ("a", None, h * a, 0.0), ("a_dd", None, 0.0, h * a_dd)]
        for lab, dVx, pa, pd in dirs:
            Js = []
            for sgn in (1.0, -1.0):
                Vx = V + sgn * h * dVx if dVx is not None else V
                c = make(prec, Vx, a + sgn * pa, a_dd + sgn * pd)
                c.rk4_steps(dt, n)
                Js.append(J(c.download_psi())[0])
                c.close()
            if dVx is not None:
                fd, ad = (Js[0] - Js[1]) / (2 * h), float(np.sum(gVn * dVx))
            elif lab == "a":
                fd, ad = (Js[0] - Js[1]) / (2 * pa), gs[0]
            else:
                fd, ad = (Js[0] - Js[1]) / (2 * pd), gs[1]
            checks.append({"dir": lab, "adjoint": ad, "fd": fd, "rel": abs(ad - fd) / abs(fd)})
            print(f"#   {lab}: adjoint {ad:.10e} fd {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e}",
                  file=sys.stderr, flush=True)
        res.update(fd_checks=checks, max_rel_vs_fd=max(x["rel"] for x in checks),
                   directional=[x["adjoint"] for x in checks])
        out[name] = res
    return {"workload": f"adjoint 128x128x128 dipolar, T = {n} RK4 steps of dt = {dt}, checkpoint every {every}",
            "box": [16.0, 16.0, 32.0], "g": g, "eps_dd": 1.3, "omega": list(omega),
            "objective": "|<target|psi(T)>|^2 wrt V (2M values), a, a_dd",
            "f64": out["f64"], "mixed": out["mixed"]}


def run_adjoint(torch, args):
    """The adjoint leg alone, as its own JSON line (16 directions by default)."""
    adj = adjoint_leg(torch, args.adjoint_steps, 45, args.adjoint_dirs)
    f64 = adj["f64"]
    line = {
        "metric": "adjoint gradients/s (fidelity wrt V, a, a_dd; 128^3 dipolar, 2000 RK4 steps, checkpoint 45)",
        "value": f64["grads_per_s"], "unit": "gradients/s", "n_gpus": 1, "steps": args.adjoint_steps,
        "warmup": 1, "ms_per_step": round(1e3 * f64["s_per_gradient"] / args.adjoint_steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Thomas-Fermi psi0 and target, harmonic trap)",
        "config": {"workload": adj["workload"]}, "f64": f64, "mixed": adj["mixed"],
    }
    print(json.dumps(line), flush=True)
    return 0


def run_config3(torch, args):
    """The two-component soft-core workload (256^3, g11 = g22 = 100, g12 = 80,
    soft-core U0 = 5, R = 1.5 on n1 + n2, dt = 5e-4): real-time RK4 steps/s in
    both forms after the autotuner, with the byte model's HBM rate and the
    mixed form's energy error against FP64 after the timed steps."""
    from fgpe_amd import FGPE_PREC_F64, FGPE_PREC_MIXED, Context, configs, roofline
    torch.cuda.set_device(0)
    w = configs.twocomp_softcore_256(nsteps=args.steps)
    g = w.grid
    psi, V = w.initial_state(), w.trap()
    kx, ky, kz = g.kaxes()
    stream = torch.cuda.current_stream()
    res = {}
    for name, prec in (("f64", FGPE_PREC_F64), ("mixed", FGPE_PREC_MIXED)):
        with Context(g.nx, g.ny, g.nz, prec) as ctx:
            ctx.set_stream(stream.cuda_stream)
            ctx.set_kspace(kx, ky, kz, g.dv)
            ctx.set_hamiltonian(ncomp=2, a=w.a, sc_U=w.sc_U, sc_R=w.sc_R, sc_dim=3)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.autotune(5)
            ctx.rk4_steps(w.dt, args.warmup, 0.0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ctx.rk4_steps(w.dt, args.steps, args.warmup * w.dt)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.steps
            obs = ctx.observables((args.warmup + args.steps) * w.dt)
            out = ctx.download_psi()
        bpp = roofline.step_bytes_per_point(name == "mixed", 2, True, g.nx, True, table=True)
        gbs = bpp * g.npoints / (ms * 1e-3) / 1e9
        res[name] = {"steps_per_s": round(1e3 / ms, 3), "ms_per_step": round(ms, 4),
                     "bytes_per_point": round(bpp, 2), "GBs": round(gbs, 1),
                     "frac_of_peak": round(gbs / roofline.HBM_PEAK_GBS, 4),
                     "energy": [float(obs[3 * c + 2]) for c in range(2)], "psi": out}
        print(f"# config 3 {name}: {1e3 / ms:.1f} steps/s, {gbs:.0f} GB/s of {bpp:.0f} B/pt", file=sys.stderr,
              flush=True)
    f, m = res["f64"], res["mixed"]
    e_rel = max(abs(m["energy"][c] - f["energy"][c]) / abs(f["energy"][c]) for c in range(2))
    l2 = float(np.linalg.norm(m["psi"] - f["psi"]) / np.linalg.norm(f["psi"]))
    for r in res.values():
        del r["psi"]
    line = {
        "metric": "RK4 steps/s at 256^3 two-component soft-core GPE (FP64 plain form; mixed form beside it)",
        "value": f["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": f["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded Gaussians offset by +-1 in x, harmonic trap)",
        "config": {"workload": "two-component soft-core real-time RK4 256x256x256", "box": [16.0, 16.0, 16.0],
                   "g": [100.0, 100.0, 80.0], "soft_core": {"U0": w.sc_U, "R": w.sc_R}, "dt": w.dt},
        "f64": f, "mixed": dict(m, speedup_vs_f64=round(m["steps_per_s"] / f["steps_per_s"], 3),
                                energy_rel_err_vs_f64=e_rel, psi_l2_rel_err_vs_f64=l2),
    }
    print(json.dumps(line), flush=True)
    return 0


def run_sweep(torch, args):
    """Batched imaginary-time ground-state sweep: every instance to
    |dE/E| < 1e-9 (heavy ball, dt = 2e-3, beta = 0.9), mixed precision,
    groups per grid shape autotuned, iterations replayed from hipGraphs of 50
    iterations: one per group, each group on its own stream (default), or one
    graph of all groups (--sweep-mode shared)."""
    from fgpe_amd import FGPE_PREC_MIXED, configs, roofline
    from fgpe_amd.sweep import Sweep
    torch.cuda.set_device(0)
    insts = configs.sweep_instances(args.sweep_n)
    t0 = time.perf_counter()
    if args.sweep_mode == "per-group":
        # one stream and one graph per shape group, each group replayed at its own pace
        streams = []

        def new_stream():
            streams.append(torch.cuda.Stream())
            return streams[-1].cuda_stream

        sw = Sweep(insts, precision=FGPE_PREC_MIXED, new_stream=new_stream)
        sync = torch.cuda.synchronize
    else:
        # every group on one stream, the groups on the branches of one graph
        side = torch.cuda.Stream()
        sw = Sweep(insts, precision=FGPE_PREC_MIXED, stream=side.cuda_stream)
        sync = side.synchronize
    setup = time.perf_counter() - t0
    print(f"# sweep: {len(insts)} instances in {len(sw.groups)} groups {sw.shapes}, set-up {setup:.1f} s",
          file=sys.stderr, flush=True)
    tune = sw.autotune(3)
    dt, beta, eps = 2e-3, 0.9, 1e-9
    iters, conv, wall, launches = sw.run(dt, beta, eps, iters_per_graph=args.sweep_iters_per_graph,
                                         max_iters=args.sweep_max_iters,
                                         recapture=args.sweep_recapture, sync=sync)
    obs = sw.energies()
    pts = np.array([i.grid.npoints for i in insts], dtype=np.float64)
    # bytes per point and iteration: x stage (heavy ball) + y/z passes, mixed precision
    bpp = []
    for i in insts:
        sc, sv, s0 = roofline.sizes(True)
        spec = 2 * roofline.channels(1, True, i.grid.nx) * sc
        xb = spec + 2 * s0 + 2 * sc + sv
        passes = (3 if i.grid.nz > 1 else 1) * roofline.pass_bytes(True, 1, True, i.grid.nx)
        bpp.append(xb + passes + (roofline.half_fraction(i.grid.nx) * sv if i.sc_U else 0.0))
    bpp = np.array(bpp)
    # the graph replays whole blocks of 50 iterations for every instance, converged ones as no-ops
    work_pts = float(np.sum(pts * iters))
    gbs = float(np.sum(pts * iters * bpp)) / wall / 1e9
    line = {
        "metric": "batched ground-state sweep: instance-iterations/s (heavy-ball imaginary time)",
        "value": round(float(iters.sum()) / wall, 1),
        "unit": "instance-iterations/s",
        "n_gpus": 1, "steps": int(iters.max()), "warmup": 0,
        "ms_per_step": round(1e3 * wall / max(1, launches * args.sweep_iters_per_graph), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 bulk, f64 state",
        "data": "synthetic (seeded Gaussian psi0, harmonic traps)",
        "config": {"workload": f"sweep of {len(insts)} dipolar / soft-core instances",
                   "shapes": [list(s) for s in sw.shapes], "dt": dt, "beta": beta, "eps": eps,
                   "iters_per_graph": args.sweep_iters_per_graph, "mode": args.sweep_mode},
        "converged": int(conv.sum()), "instances": len(insts),
        "iterations": {"min": int(iters.min()), "max": int(iters.max()), "total": int(iters.sum())},
        "wall_s": round(wall, 3), "graph_launches": launches, "graph_captures": sw.captures,
        "recapture": args.sweep_recapture,
        "point_iterations_per_s": round(work_pts / wall, 1),
        "hbm_GBs_model": round(gbs, 1),
        "autotune": {"x".join(map(str, k)): {c: [round(x, 4) for x in v] for c, v in t.items()}
                     for k, t in tune.items()},
        "block_order": {"x".join(map(str, shp)): sw.ctxs[m[0]].order() for shp, m in zip(sw.shapes, sw.members)},
        "energies": [round(float(e), 10) for e in obs[:, 2]],
    }
    if hasattr(sw, "group_finish_s"):
        per = {}
        for shp, m, fin in zip(sw.shapes, sw.members, sw.group_finish_s):
            per["x".join(map(str, shp))] = {"instances": len(m), "iterations_max": int(iters[m].max()),
                                            "iterations_total": int(iters[m].sum()), "finish_s": round(fin, 4)}
        line["groups"] = per
    sw.close()
    print(json.dumps(line), flush=True)
    return 0


def run_sweep_kernels(torch, args):
    """Per-kernel-class HBM throughput of the sweep's batched launches: each
    shape group of the 64-instance sweep runs 20 heavy-ball iterations as
    eager batched launches with every instance active (eps = 0), bracketed
    by HIP events per launch; the bytes per launch are the byte model's over
    the group's instances (fgpe_amd/roofline.py)."""
    from fgpe_amd import FGPE_PREC_MIXED, configs, roofline
    from fgpe_amd.sweep import Sweep
    torch.cuda.set_device(0)
    insts = configs.sweep_instances(args.sweep_n)
    side = torch.cuda.Stream()
    sw = Sweep(insts, precision=FGPE_PREC_MIXED, stream=side.cuda_stream)
    sw.autotune(3)
    sc, sv, s0 = roofline.sizes(True)
    rows = []
    for shape, members, group in zip(sw.shapes, sw.members, sw.groups):
        nx, ny, nz = shape
        c0 = sw.ctxs[members[0]]
        group.imag_steps(2e-3, 2, beta=0.9, npart=1.0, eps=0.0)   # warm-up
        side.synchronize()
        c0.timing_enable(True)
        iters = 20
        group.imag_steps(2e-3, iters, beta=0.9, npart=1.0, eps=0.0)
        kt = c0.timing_read()
        c0.timing_enable(False)
        pts = nx * ny * nz
        spec = 2 * roofline.channels(1, True, nx) * sc
        bpp = {"xstage": sum(spec + 2 * s0 + 2 * sc + sv for _ in members),
               "strided": sum(roofline.pass_bytes(True, 1, True, nx) for _ in members),
               "last": sum(roofline.pass_bytes(True, 1, True, nx, table=bool(insts[i].sc_U)) for i in members)}
        for cls in ("xstage", "strided", "last"):
            ms, n = kt[cls]
            if n == 0:
                continue
            per = ms / n
            gbs = bpp[cls] * pts / (per * 1e-3) / 1e9
            rows.append({"shape": list(shape), "instances": len(members), "kernel": cls, "launches": n,
                         "avg_us": round(1e3 * per, 2), "bytes_per_launch": round(bpp[cls] * pts),
                         "GBs": round(gbs, 1), "frac_of_peak": round(gbs / roofline.HBM_PEAK_GBS, 4)})
            print(f"# {shape} x{len(members)} {cls}: {1e3 * per:.1f} us, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
        ms, n = kt["other"]
        if n:
            rows.append({"shape": list(shape), "instances": len(members), "kernel": "finalize", "launches": n,
                         "avg_us": round(1e3 * ms / n, 2)})
    sw.close()
    print(json.dumps({"metric": "sweep batched launches, per kernel class: HBM GB/s of the byte model",
                      "unit": "GB/s", "n_gpus": 1, "dtype": "f32 bulk, f64 state", "rows": rows}), flush=True)
    return 0


def run_init(torch, args):
    """Device initialisation of the state and trap at the RK4 workload's grid:
    harmonic psi with complex noise, normalisation, a vortex imprint, the
    harmonic trap plus a cosine seed, committed to the Hamiltonian.  One
    "step" is the whole sequence (six launch groups, one host read of the
    norm).  The oracle's C restatement of the same loops is the CPU
    baseline."""
    from fgpe_amd import FGPE_PREC_F64, Context, Geom
    from oracle import oracle as orc
    torch.cuda.set_device(0)
    nx, ny, nz = (int(v) for v in args.grid.split(","))
    box = (8.0, 8.0, 16.0)
    dx = [2 * b / n for b, n in zip(box, (nx, ny, nz))]
    ax = (1.0, 1.0, 2.0)
    g = Geom.of(box, dx)
    npts = nx * ny * nz

    def seq(ctx):
        ctx.init_psi(g, 0, ax, rand=(0.01, 0.01), seed=367)
        ctx.psi_normalize(g, 1.0e5)
        ctx.psi_imprint_vortex_tube(g, 1.0)
        ctx.vext_init(g, 0, ax)
        ctx.vext_add_cos(g, 0.01, 2.0)
        ctx.vext_commit()

    with Context(nx, ny, nz, precision=FGPE_PREC_F64) as ctx:
        for _ in range(args.warmup):
            seq(ctx)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            seq(ctx)
        ctx.sync()
        wall = (time.perf_counter() - t0) / args.steps
        got = ctx.download_psi()
    # HBM bytes per point of the sequence: init writes psi (16); normalise
    # reads it, then reads and writes it (48); the vortex reads and writes
    # it (32); the trap is written (8), read and written by the cosine seed
    # (16) and read once more by the commit, which writes the bulk trap (16)
    bpp = 16 + 48 + 32 + 8 + 16 + 16
    achieved = bpp * npts / wall / 1e9
    cpu = None
    if not args.no_cpu:
        og = orc.geom((nx, ny, nz), box, dx)
        orc.set_threads(args.cpu_threads)
        t0 = time.perf_counter()
        ref = orc.psi_init(og, 0, ax, rand=(0.01, 0.01), seed=367)
        ref, _ = orc.psi_normalize(og, ref, 1.0e5)
        ref = orc.psi_vortex(og, ref, 1.0)
        v = orc.vext_add_cos(og, orc.vext_init(og, 0, ax), 0.01, 2.0)
        ct = time.perf_counter() - t0
        err = float(np.abs(got.reshape(-1) - ref).max() / np.abs(ref).max())
        cpu = {"value": round(1.0 / ct, 4), "unit": "initialisations/s", "cores": orc.num_threads(),
               "kind": "port", "sample": f"one full {nx}x{ny}x{nz} sequence ({ct:.2f} s)",
               "max_rel_diff_vs_device": err}
        del v
    line = {
        "metric": "device initialisations/s of psi and Vext (harmonic psi + noise, normalise, vortex, "
                  "harmonic trap + cosine seed)",
        "value": round(1.0 / wall, 2), "unit": "initialisations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (generated on the device)",
        "config": {"workload": f"initial state and trap {nx}x{ny}x{nz}", "box": list(box), "ax": list(ax)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "note": f"whole sequence, {bpp} B/point over the wall time including launch gaps "
                             "and one device-to-host read"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="256,256,256")
    ap.add_argument("--cpu-steps", type=int, default=20, help="timed oracle steps on --cpu-threads threads")
    ap.add_argument("--cpu-warmup", type=int, default=2)
    ap.add_argument("--cpu-steps-1t", type=int, default=2, help="timed oracle steps on one thread")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="oracle threads (16 is a GPU box's CPU share; capped at the host's count)")
    ap.add_argument("--no-adjoint", action="store_true", help="leave the adjoint leg out of the default line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", action="store_true",
                    help="run the batched ground-state sweep (64 mixed-shape dipolar / soft-core "
                         "instances, one hipGraph) instead of the 256^3 RK4 benchmark")
    ap.add_argument("--adjoint", action="store_true",
                    help="time adjoint gradients (128^3 dipolar, 2000 RK4 steps, checkpoint every 45) "
                         "in both forms, with central finite-difference checks")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-dirs", type=int, default=0,
                    help="seeded smooth V directions checked by central differences (default: 4 in the "
                         "default line, 16 with --adjoint)")
    ap.add_argument("--config3", action="store_true",
                    help="time the two-component soft-core workload (256^3) in both forms instead")
    ap.add_argument("--init", action="store_true",
                    help="time the device initialisers (state and trap) at --grid instead")
    ap.add_argument("--sweep-kernels", action="store_true",
                    help="time the sweep's batched launches per kernel class and shape group (HBM GB/s)")
    ap.add_argument("--sweep-n", type=int, default=64)
    ap.add_argument("--sweep-mode", choices=("per-group", "shared"), default="per-group",
                    help="sweep replays: one graph per shape group on its own stream, or one graph of all groups")
    ap.add_argument("--sweep-max-iters", type=int, default=20000)
    ap.add_argument("--sweep-iters-per-graph", type=int, default=50)
    ap.add_argument("--sweep-recapture", type=float, default=0.75,
                    help="re-capture the graph over the unconverged instances once this share of the "
                         "instances in it has converged")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 launches), the known-size "
                         "reference that scales FETCH_SIZE / WRITE_SIZE counter readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path is "
              "single-GPU by design (no sharded form is built)", file=sys.stderr)
        return 2

    import torch
    if args.copy_calib > 0:
        n = args.copy_calib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_each_way": n * 8, "launches": 3}), flush=True)
        return 0
    from fgpe_amd import FGPE_PREC_F64, FGPE_PREC_MIXED, configs, roofline
    if args.sweep_kernels:
        return run_sweep_kernels(torch, args)
    if args.sweep:
        return run_sweep(torch, args)
    if args.adjoint:
        args.adjoint_dirs = args.adjoint_dirs or 16
        return run_adjoint(torch, args)
    if args.init:
        return run_init(torch, args)
    if args.config3:
        return run_config3(torch, args)

    nx, ny, nz = (int(v) for v in args.grid.split(","))
    w = configs.dipolar_real(nx, ny, nz, nsteps=args.steps)
    psi, V = w.initial_state(), w.trap()
    torch.cuda.set_device(0)
    print(f"# workload {w.name}: {nx}x{ny}x{nz}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)

    res = {}
    for name, prec in (("f64", FGPE_PREC_F64), ("mixed", FGPE_PREC_MIXED)):
        ms, kt, obs, out, tune = run_form(torch, w, psi, V, prec, args.steps, args.warmup)
        res[name] = dict(ms=ms, kt=kt, obs=obs, psi=out, tune=tune)
        print(f"# {name}: {ms:.3f} ms/step, {1000.0 / ms:.2f} steps/s, kernels {kt}, tune {tune}",
              file=sys.stderr, flush=True)

    N = nx * ny * nz
    f64 = res["f64"]
    mixed = res["mixed"]
    steps_s = 1000.0 / f64["ms"]
    # dominant kernel class of the FP64 form
    kt = {k: v for k, v in f64["kt"].items() if v[1] > 0}
    dom = max(kt, key=lambda k: kt[k][0])
    avg_ms = kt[dom][0] / kt[dom][1]
    if dom == "xstage":
        per_pt = sum(roofline.xstage_bytes(s, False, 1, True, nx) for s in (1, 2, 3, 4)) / 4.0
    else:
        per_pt = roofline.pass_bytes(False, 1, True, nx)
    achieved = per_pt * N / (avg_ms * 1e-3) / 1e9
    traffic = None
    if dom == "xstage" and (nx, ny, nz) == (256, 256, 256):
        v = f64["tune"]["variant"]["xstage"]
        # the 256^3 working set exceeds the 256 MB memory-side cache, so the
        # x stage runs its non-temporal form (last template argument true)
        traffic = pmc_traffic(f"fgpe::k_xstage<double, 256, {v}, 1, true, 1, true, 0>")
    step_gbs = roofline.step_bytes_per_point(False, 1, True, nx) * N * steps_s / 1e9
    mixed_steps = 1000.0 / mixed["ms"]
    l2 = float(np.sqrt(np.sum(np.abs(mixed["psi"] - f64["psi"]) ** 2) / np.sum(np.abs(f64["psi"]) ** 2)))
    de = float(abs(mixed["obs"][2] - f64["obs"][2]) / abs(f64["obs"][2]))

    line = {
        "metric": "RK4 steps/s at 256^3 dipolar GPE (FP64 plain form; mixed form beside it)",
        "value": round(steps_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Thomas-Fermi psi0 + noise, harmonic trap)",
        "config": {"workload": f"dipolar real-time RK4 {nx}x{ny}x{nz}", "grid": [nx, ny, nz],
                   "box": [16.0, 16.0, 32.0], "g": 200.0, "eps_dd": 1.3, "omega": [1, 1, 0.25],
                   "dt": w.dt},
        "mixed": {"steps_per_s": round(mixed_steps, 3), "ms_per_step": round(mixed["ms"], 4),
                  "speedup_vs_f64": round(mixed_steps / steps_s, 3), "l2_psi_err_vs_f64": l2,
                  "rel_energy_err_vs_f64": de, "steps_compared": args.warmup + 2 * args.steps},
        "step_traffic": {"bytes_per_point_f64": round(roofline.step_bytes_per_point(False, 1, True, nx), 2),
                         "bytes_per_point_mixed": round(roofline.step_bytes_per_point(True, 1, True, nx), 2),
                         "f64_GBs": round(step_gbs, 1),
                         "f64_frac_of_peak": round(step_gbs / roofline.HBM_PEAK_GBS, 4)},
        "autotune": {"f64": f64["tune"], "mixed": mixed["tune"]},
        "kernels_f64_ms_avg": {k: round(v[0] / v[1], 4) for k, v in kt.items()},
        "kernels_mixed_ms_avg": {k: round(v[0] / v[1], 4) for k, v in mixed["kt"].items() if v[1] > 0},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": roofline.HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / roofline.HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": PMC_SUMMARY if traffic else None,
                     "algorithmic_bytes_per_launch": round(per_pt * N),
                     "bytes_per_point": per_pt, "avg_launch_ms": round(avg_ms, 4)},
        "cpu_baseline": None,
    }
    if not args.no_adjoint:
        t0 = time.perf_counter()
        line["adjoint"] = adjoint_leg(torch, args.adjoint_steps, 45, args.adjoint_dirs or 4)
        line["adjoint"]["leg_wall_s"] = round(time.perf_counter() - t0, 1)
    if not args.no_cpu and args.cpu_steps > 0:
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        sps, used, sps1, wall = cpu_baseline(w, psi, V, args.cpu_steps, args.cpu_warmup, threads, args.cpu_steps_1t)
        line["cpu_baseline"] = {
            "value": round(sps, 5), "unit": "steps/s", "cores": used, "kind": "port",
            "sample": f"{args.cpu_steps} FP64 RK4 steps (after {args.cpu_warmup} untimed) of the same "
                      f"{nx}x{ny}x{nz} dipolar workload, oracle/fgpe_oracle.c (its own FFT; no FFTW here), "
                      f"{used} OpenMP threads",
            "one_thread_steps_per_s": round(sps1, 5) if sps1 else None,
            "one_thread_steps": args.cpu_steps_1t, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "wall_s": round(wall, 1),
            "gpu_f64_speedup": round(steps_s / sps, 1)}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
