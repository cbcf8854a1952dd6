"""Steps captured into HIP graphs and replayed (use_graph = 1) give results
bit-identical to the eagerly launched steps, replay after replay, with eager
readbacks between the replays."""
import numpy as np
import pytest

from graphics_amd import scenes, sphgpu as sg

pytestmark = pytest.mark.gpu


def _trajectory(sc, use_graph, steps, every=3, net=None, **kw):
    sim = scenes.make_simulation(sc, use_graph=use_graph, **kw)
    if net is not None:
        sim.set_mlp_params(net)
    out = []
    for k in range(steps):
        st = sim.step()
        if k % every == 0:
            out.append((k, sim.get(sg.F_POS).tobytes(), sim.get(sg.F_VEL).tobytes(),
                        st["iter_density"], st["density_err"], st["num_fluid"]))
    sim.close()
    return out


def _net():
    """A network with a non-zero output layer (random, scaled down), so that
    the learned guess differs from the previous stiffness."""
    m = sg.Mlp(64, seed=3)
    p = m.get_params()
    m.close()
    p[-65:] *= 0.05
    return p


@pytest.mark.parametrize("case", ["wcsph", "iisph_batched", "dfsph_plain", "dfsph_previous", "dfsph_learned"])
def test_graph_replay_matches_eager(gpu, case):
    """Whole captured steps (WCSPH, IISPH: one graph; DFSPH: one per solver
    loop, with a plain, previous-stiffness or learned start of the
    divergence loop) against eager launches."""
    kw, net = {}, None
    if case == "wcsph":
        sc = scenes.small_dam(sg.WCSPH, n=(10, 12, 10), box=(0.5, 0.4, 0.35))
        steps = 40
    elif case == "iisph_batched":
        sc = scenes.iisph_batched_c4(scenes=3, nmin=300, nmax=2000, seed=11)
        steps = 24
    else:
        sc = scenes.small_dam(sg.DFSPH, n=(10, 12, 10), box=(0.5, 0.4, 0.35))
        steps = 24
        kw = dict(warm_start={"dfsph_plain": sg.WARM_NONE, "dfsph_previous": sg.WARM_PREVIOUS,
                              "dfsph_learned": sg.WARM_LEARNED}[case])
        if case == "dfsph_learned":
            net = _net()
    eager = _trajectory(sc, 0, steps, net=net, **kw)
    graph = _trajectory(sc, 1, steps, net=net, **kw)
    assert len(eager) == len(graph)
    for a, b in zip(eager, graph):
        assert a == b, f"step {a[0]} differs"



def test_autotuner_choices_reported(gpu):
    sc = scenes.small_dam(sg.DFSPH, n=(10, 10, 10))
    sim = scenes.make_simulation(sc)
    with pytest.raises(sg.SphError):
        sim.tuning()                     # runs before the first step
    sim.step()
    t = sim.tuning()
    assert t["list_stage_groups"] in (6, 8)
    assert all(v > 0 for v in t["ms"].values())
    # the LDS window of each DFSPH iteration pass was timed both ways, and
    # the faster form kept
    w = sim.lds_windows()
    dfsph = sg.WIN_DENSITY | sg.WIN_DIVERGENCE | sg.WIN_VELOCITY
    assert w["mask"] & ~dfsph == sg.WIN_SUMDIJ | sg.WIN_PRESSURE
    for name, bit in (("density", sg.WIN_DENSITY), ("divergence", sg.WIN_DIVERGENCE),
                      ("velocity", sg.WIN_VELOCITY)):
        gathers, window = w["ms"][name]
        assert gathers > 0 and window > 0
        assert bool(w["mask"] & bit) == (window <= gathers)
        # a gathering pass runs at its register-limited occupancy or capped
        # at 6, 5, 4 or 3 blocks per CU; a windowed pass reports no cap
        assert w["blocks_per_cu"][name] in ((0,) if w["mask"] & bit else (0, 3, 4, 5, 6))
    assert w["ms"]["sumdij"] is None and w["ms"]["pressure"] is None
    sim.close()


@pytest.mark.parametrize("use_graph", [0, 1])
def test_steps_without_statistics_match(gpu, use_graph):
    """Steps that read nothing back (and do not wait) give the same state."""
    sc = scenes.small_dam(sg.WCSPH, n=(8, 10, 8), box=(0.4, 0.3, 0.3))
    a = scenes.make_simulation(sc, use_graph=use_graph)
    b = scenes.make_simulation(sc, use_graph=use_graph)
    for k in range(30):
        a.step()
        assert b.step(stats=(k % 7 == 6)) is None or k % 7 == 6
    st = b.step()
    a.step()
    assert a.get(sg.F_POS).tobytes() == b.get(sg.F_POS).tobytes()
    assert st["num_fluid"] == a.num_fluid()
    a.close()
    b.close()


@pytest.mark.parametrize("shape", ["cube", "long"])
def test_tuning_choices_are_bit_identical(gpu, shape):
    """Every autotuner choice (tile mapping, list staging depth, plain or
    streaming list loads) gives bit-identical steps, on a scene with 32-bit
    lists (cube) and one with packed lists (long tank)."""
    kw = dict(n=(12, 12, 12)) if shape == "cube" else dict(n=(96, 6, 8), box=(2.0, 0.3, 0.3))
    runs = []
    for fix in [(0, 6, 0), (1, 8, 1), (1, 6, 1), (0, 8, 0)]:
        sc = scenes.small_dam(sg.DFSPH, **kw)
        sim = scenes.make_simulation(sc, warm_start=sg.WARM_PREVIOUS)
        sim.set_tuning(*fix)
        st = [sim.step() for _ in range(6)]
        t = sim.tuning()
        assert (int(t["tile_map_xcd"]), t["list_stage_groups"], int(sim.list_form()["streamed"])) == fix
        runs.append((sim.get(sg.F_POS).tobytes(), sim.get(sg.F_VEL).tobytes(),
                     [s["iter_density"] for s in st], [s["iter_divergence"] for s in st]))
        sim.close()
    assert all(r == runs[0] for r in runs[1:])


@pytest.mark.parametrize("solver,warm", [(sg.DFSPH, sg.WARM_PREVIOUS), (sg.DFSPH, sg.WARM_LEARNED), (sg.IISPH, 0)])
def test_lds_windows_are_bit_identical(gpu, solver, warm):
    """Solver-loop passes reading neighbours from LDS windows give the same
    bits as per-lane gathers, on a tank whose tiles have packed lists, with
    the previous-stiffness and the learned DFSPH start (whose divergence
    solve starts from the projected iteration) and for IISPH."""
    runs = []
    for on in (1, 0):
        sc = scenes.small_dam(solver, n=(96, 6, 8), box=(2.0, 0.3, 0.3))
        sim = scenes.make_simulation(sc, warm_start=warm)
        sim.set_lds_windows(on)
        st = [sim.step() for _ in range(6)]
        assert sim.list_form()["packed_tiles"] > 0
        runs.append((sim.get(sg.F_POS).tobytes(), sim.get(sg.F_VEL).tobytes(),
                     [s["iter_density"] for s in st], [s["iter_divergence"] for s in st]))
        sim.close()
    assert runs[0] == runs[1]
