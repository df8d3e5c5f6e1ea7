"""Batched groups and captured graphs (the sweep path): a group of same-shape
contexts must reproduce, bit for bit, what each context computes on its own;
graph replays must reproduce the eager batched iterations; and a small mixed
precision sweep run to convergence must land on the FP64 oracle's ground
state energies (relative error bound written at the assertion)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.cases import Case

pytestmark = pytest.mark.gpu

DT, BETA = 2e-3, 0.9

# one-component instances of different kinds, couplings and traps
INSTANCES = [
    dict(h=dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0)), omega=(1.0, 1.0, 1.3), seed=21),
    dict(h=dict(a=(0.08,), sc_U=2.0, sc_R=1.2, sc_dim=3), omega=(1.2, 1.0, 1.0), seed=22),
    dict(h=dict(a=(0.03,), a_dd=0.02, dip=(1.0, 0.0, 0.0)), omega=(1.0, 1.1, 0.9), seed=23),
]


def make_ctx(case, inst, precision, stream=None):
    from fgpe_amd import configs
    c = case.ctx(precision)
    if stream is not None:
        c.set_stream(stream)
    h = dict(inst["h"])
    if case.spec.nz == 1 and h.get("sc_U"):
        h["sc_dim"] = 2
    c.set_hamiltonian(**h)
    c.upload_psi(configs.gaussian_noisy(case.spec, inst["seed"], width=1.2))
    c.upload_vext(configs.harmonic_trap(case.spec, inst["omega"]))
    return c


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 64, 1), (32, 16, 256), (32, 256, 1)])
def test_group_equals_single_contexts(torch_cuda, precision, shape):
    """The 256-long last axes run the FP64 last-axis pass's split-exchange
    plan (plan 0, the default) in batched launches."""
    from fgpe_amd import FGPE_IMAG_HEAVY_BALL, Group
    side = torch_cuda.cuda.Stream()
    case = Case(*shape, L=tuple(12.0 if n <= 64 else 24.0 for n in shape))
    singles = [make_ctx(case, inst, precision) for inst in INSTANCES]
    batch = [make_ctx(case, inst, precision, side.cuda_stream) for inst in INSTANCES]
    group = Group(batch)
    try:
        n = 25
        for c in singles:
            assert c.imag_steps(DT, n, scheme=FGPE_IMAG_HEAVY_BALL, beta=BETA, npart=1.0, eps=0.0) == n
        group.imag_steps(DT, n, beta=BETA, npart=1.0, eps=0.0)
        iters, conv = group.imag_finish()
        assert list(iters) == [n] * len(INSTANCES), iters
        assert not conv.any()
        for i, (a, b) in enumerate(zip(singles, batch)):
            x, y = a.download_psi(), b.download_psi()
            err = float(np.abs(x - y).max())
            print(f"instance {i}: max |batched - single| = {err:.3e}")
            assert err == 0.0, (i, err)
    finally:
        group.close()
        for c in singles + batch:
            c.close()


def test_graph_replays_match_eager_batched(torch_cuda):
    """Two shapes (3D and 2D), 10 iterations per replay, three replays with
    downloads in between; every replay equals the eager batched run."""
    from fgpe_amd import Graph, Group
    torch = torch_cuda
    side = torch.cuda.Stream()
    stream = side.cuda_stream
    cases = [Case(32, 32, 32, L=(12.0, 12.0, 12.0)), Case(64, 64, 1, L=(12.0, 12.0, 12.0))]
    eager = [[make_ctx(cs, inst, 1, stream) for inst in INSTANCES[:2]] for cs in cases]
    graphed = [[make_ctx(cs, inst, 1, stream) for inst in INSTANCES[:2]] for cs in cases]
    ge = [Group(cs) for cs in eager]
    gg = [Group(cs) for cs in graphed]
    graph = Graph(gg, DT, 10, beta=BETA, npart=1.0, eps=0.0)
    try:
        for r in range(3):
            for g in ge:
                g.imag_steps(DT, 10, beta=BETA, npart=1.0, eps=0.0)
            graph.launch()
            side.synchronize()
            for k in range(len(cases)):
                _, ce = ge[k].imag_status()
                it, cg = gg[k].imag_status()
                assert list(it) == [10 * (r + 1)] * 2, (r, k, it)
                for a, b in zip(eager[k], graphed[k]):
                    err = float(np.abs(a.download_psi() - b.download_psi()).max())
                    print(f"replay {r} group {k}: max |graph - eager| = {err:.3e}")
                    assert err == 0.0, (r, k, err)
    finally:
        graph.close()
        for g in ge + gg:
            g.close()
        for cs in eager + graphed:
            for c in cs:
                c.close()


def test_small_sweep_converges_to_oracle_ground_states(torch_cuda):
    """Mixed precision (FP32 bulk, FP64 state and reductions), eps = 1e-9 on
    |dE/E|, graph replays of 25 iterations; energies within 1e-6 relative of
    the FP64 oracle's heavy-ball ground states."""
    from fgpe_amd import configs
    from fgpe_amd.sweep import Sweep
    from tests.test_gpu_parity import oracle_ham
    torch = torch_cuda
    side = torch.cuda.Stream()
    eps = 1e-9
    insts = []
    for k, (shape, inst) in enumerate([((32, 32, 32), INSTANCES[0]), ((32, 32, 32), INSTANCES[1]),
                                       ((64, 64, 1), INSTANCES[2])]):
        spec = configs.GridSpec(shape[0], shape[1], shape[2], 12.0, 12.0, 12.0 if shape[2] > 1 else 1.0)
        h = inst["h"]
        si = configs.SweepInstance(k, spec, h["a"][0], inst["omega"], a_dd=h.get("a_dd", 0.0),
                                   dip=h.get("dip", (0.0, 0.0, 1.0)), sc_U=h.get("sc_U", 0.0),
                                   sc_R=h.get("sc_R", 0.0), sc_dim=h.get("sc_dim", 3) if shape[2] > 1 else 2,
                                   seed=inst["seed"])
        si.initial_state = (lambda s=spec, sd=inst["seed"]: configs.gaussian_noisy(s, sd, width=1.2))
        insts.append(si)
    sw = Sweep(insts, precision=1, stream=side.cuda_stream)
    try:
        iters, conv, wall, launches = sw.run(DT, BETA, eps, iters_per_graph=25, max_iters=6000,
                                             sync=side.synchronize)
        print("iterations", iters, "converged", conv, f"{wall:.3f} s in {launches} replays")
        assert conv.all(), (iters, conv)
        obs = sw.energies()
    finally:
        sw.close()
    for k, si in enumerate(insts):
        c = Case(si.grid.nx, si.grid.ny, si.grid.nz, L=(si.grid.lx, si.grid.ly, si.grid.lz))
        h = dict(a=(si.a,), a_dd=si.a_dd, dip=si.dip, sc_U=si.sc_U, sc_R=si.sc_R, sc_dim=si.sc_dim)
        V = si.trap()
        psi0 = si.initial_state()
        ref, nref, href = orc.imag_heavy_ball(c.grid, oracle_ham(c, h), psi0, V, DT, BETA, 1.0, eps, 6000)
        e_ref = href[-1, 2]
        rel = abs(obs[k, 2] - e_ref) / abs(e_ref)
        print(f"instance {k} ({si.kind}, {si.grid.nx}x{si.grid.ny}x{si.grid.nz}): E gpu {obs[k, 2]:.12f} "
              f"oracle {e_ref:.12f} rel {rel:.2e}; iterations gpu {iters[k]} oracle {nref}")
        assert nref < 6000
        assert rel < 1e-6, (k, rel)


@pytest.mark.parametrize("precision", [0, 1])
def test_sweep_shapes_against_oracle(torch_cuda, precision):
    """The sweep configuration's own grids (96^3, 64 x 64 x 192, 2D 512^2),
    two instances each (dipolar and soft-core), run as batched groups on the
    branches of one captured graph: 40 heavy-ball iterations (two replays of
    20) against the oracle's 40 iterations of each instance, the FP64 form
    within 1e-11 on psi and 1e-11 on the energy, the mixed form within 1e-6
    on the energy."""
    from fgpe_amd import configs
    from fgpe_amd.sweep import Sweep
    from tests.test_gpu_parity import oracle_ham
    torch = torch_cuda
    side = torch.cuda.Stream()
    insts = configs.sweep_instances(6, seed=7, shapes=((96, 96, 96), (64, 64, 192), (512, 512, 1)))
    sw = Sweep(insts, precision=precision, stream=side.cuda_stream)
    try:
        iters, conv, wall, launches = sw.run(DT, BETA, 0.0, iters_per_graph=20, max_iters=40, sync=side.synchronize)
        assert list(iters) == [40] * len(insts) and not conv.any(), (iters, conv)
        states = [c.download_psi() for c in sw.ctxs]
        obs = sw.energies()
    finally:
        sw.close()
    orc.set_threads(16)
    for k, si in enumerate(insts):
        c = Case(si.grid.nx, si.grid.ny, si.grid.nz, L=(si.grid.lx, si.grid.ly, si.grid.lz))
        h = dict(a=(si.a,), a_dd=si.a_dd, dip=si.dip, sc_U=si.sc_U, sc_R=si.sc_R, sc_dim=si.sc_dim)
        ham, V = oracle_ham(c, h), si.trap()
        ref, nref, href = orc.imag_heavy_ball(c.grid, ham, si.initial_state(), V, DT, BETA, 1.0, 0.0, 40)
        # the energy of the final state (the history's E lags one iteration)
        hmu, hen = orc.hpsi(c.grid, ham, 0.0, ref, V)
        e_ref = orc.obs2(c.grid, orc.norm(c.grid, ref), ref, hmu, hen)[1]
        e_psi = float(np.abs(states[k] - ref).max() / np.abs(ref).max())
        e_en = abs(obs[k, 2] - e_ref) / abs(e_ref)
        print(f"precision {precision} instance {k} ({si.kind}, {si.grid.nx}x{si.grid.ny}x{si.grid.nz}): "
              f"psi {e_psi:.2e}, E {e_en:.2e}", flush=True)
        if precision == 0:
            assert e_psi < 1e-11 and e_en < 1e-11, (k, e_psi, e_en)
        else:
            assert e_en < 1e-6, (k, e_en)


def test_sweep_per_group_streams_match_one_graph(torch_cuda):
    """The sweep with one stream and one graph per shape group (each group
    replayed and recaptured at its own pace) gives every instance the same
    iteration count and the same ground state as one graph of all groups."""
    from fgpe_amd import configs
    from fgpe_amd.sweep import Sweep
    torch = torch_cuda
    insts = configs.sweep_instances(6, seed=11, shapes=((32, 32, 32), (64, 32, 16), (128, 128, 1)))
    side = torch.cuda.Stream()
    runs = []
    for mode in ("shared", "per-group"):
        keep = []

        def new_stream():
            keep.append(torch.cuda.Stream())
            return keep[-1].cuda_stream

        sw = (Sweep(insts, precision=1, stream=side.cuda_stream) if mode == "shared"
              else Sweep(insts, precision=1, new_stream=new_stream))
        try:
            iters, conv, wall, launches = sw.run(DT, BETA, 1e-9, iters_per_graph=20, max_iters=4000,
                                                 sync=torch.cuda.synchronize)
            runs.append((iters, conv, sw.energies(), [c.download_psi() for c in sw.ctxs]))
            print(mode, iters, conv, f"{wall:.3f} s, {launches} replays", flush=True)
        finally:
            sw.close()
    (i0, c0, e0, p0), (i1, c1, e1, p1) = runs
    assert c0.all() and c1.all()
    assert list(i0) == list(i1)
    assert np.array_equal(e0, e1)
    for a, b in zip(p0, p1):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("order", [5, 10, 15])
def test_group_block_order_is_bitwise_neutral(torch_cuda, order):
    """Batched heavy-ball iterations with passes run reversed (tiles and
    instances, the order the group autotuner may pick) give the same states,
    bit for bit, as the forward order."""
    from fgpe_amd import Group, configs
    insts = configs.sweep_instances(3, seed=13, shapes=((32, 32, 32),))
    out = []
    for o in (0, order):
        side = torch_cuda.cuda.Stream()
        ctxs = []
        for inst in insts:
            g = inst.grid
            c = Case(g.nx, g.ny, g.nz, L=(g.lx, g.ly, g.lz)).ctx(1)
            c.set_stream(side.cuda_stream)
            c.set_hamiltonian(a=(inst.a,), a_dd=inst.a_dd, dip=inst.dip, sc_U=inst.sc_U, sc_R=inst.sc_R,
                              sc_dim=inst.sc_dim)
            c.upload_psi(inst.initial_state())
            c.upload_vext(inst.trap())
            c.set_order(o)
            ctxs.append(c)
        grp = Group(ctxs)
        try:
            grp.imag_steps(DT, 12, beta=BETA, npart=1.0, eps=0.0)
            grp.imag_finish()
            out.append([c.download_psi() for c in ctxs])
        finally:
            grp.close()
            for c in ctxs:
                c.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
