"""The device-side time loop (dt, t_end clamp, dt < 1e-14 stop, time and step
count on the GPU; chunks of steps captured in a HIP graph and replayed) against
the host-driven loop of src/sdader_simulator.py:394-413.

Both loops launch the same kernels with the same dt arithmetic, so the bar is
bit-identical states, times and step counts.
"""
import numpy as np
import pytest
import torch

from fixtures import golden_names, load, make_ic, noise_of, sim_kwargs

pytestmark = pytest.mark.gpu

EVOLUTION = [n for n in golden_names() if "final_U_sp" in load(n)]


def make(fx, **kw):
    from spd_amd.simulator import SDADER_Simulator
    cfg = fx["cfg"]
    s = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                         **sim_kwargs(cfg), **kw)
    nz = noise_of(cfg)
    if nz is not None:
        s.perturb_initial_condition(nz)
    return s


def same_run(a, b):
    assert a.n_step == b.n_step
    assert a.time == b.time
    assert a.dm.dt == b.dm.dt
    np.testing.assert_array_equal(a.dm.U_sp, b.dm.U_sp)
    np.testing.assert_array_equal(a.dm.W_cv, b.dm.W_cv)


@pytest.mark.parametrize("graph_steps", [5, 0])
@pytest.mark.parametrize("name", EVOLUTION)
def test_time_evolution_device_loop(name, graph_steps):
    print(f"[device loop] {name} graph_steps={graph_steps}", flush=True)
    fx = load(name)
    t_end = fx["cfg"]["t_end"]
    host = make(fx, device_loop=False)
    host.perform_time_evolution(t_end)
    dev = make(fx, device_loop=True, graph_steps=graph_steps)
    dev.perform_time_evolution(t_end)
    same_run(host, dev)
    assert dev.n_step == int(fx["final_n_step"])


@pytest.mark.parametrize("name", ["kh16_p3", "quad10_p3_mixed", "sod_p1_llf", "step3d_p2"])
def test_iterations_device_loop(name):
    """perform_iterations(7) with 3-step graphs: two replays and a 1-step tail of
    direct launches; then a second call continues from the first."""
    print(f"[device loop iterations] {name}", flush=True)
    fx = load(name)
    host = make(fx, device_loop=False)
    dev = make(fx, device_loop=True, graph_steps=3)
    for n in (7, 4):
        host.perform_iterations(n)
        dev.perform_iterations(n)
        same_run(host, dev)


def test_graph_replays_each_checked():
    """One captured 4-step graph replayed five times, with eager work (dt
    reduction, run-state reset, status read-back) between replays; every replay
    is compared with four host-loop steps."""
    fx = load("kh16_p3")
    host = make(fx, device_loop=False)
    dev = make(fx, device_loop=True, graph_steps=4)
    host.init_sim()
    dev.init_sim()
    graphs = set()
    for r in range(5):
        print(f"[graph replay] {r}", flush=True)
        for _ in range(4):
            host.compute_dt()
            host._step(snapshot=False)
        dev._run_on_device(0.0, 4, 0)
        graphs.add(id(dev._graph[1]))
        torch.cuda.synchronize()
        assert dev.n_step == host.n_step and dev.time == host.time
        assert torch.equal(dev.dm.device_value("U_sp"), host.dm.device_value("U_sp")), r
    assert len(graphs) == 1
    host.end_sim()
    dev.end_sim()
    same_run(host, dev)


def test_dt_floor_stop():
    """A run whose clamped dt falls below 1e-14 stops without the step, as the
    reference's loop does; the reported dt is the clamped one."""
    fx = load("sod_p3")
    t_end = fx["cfg"]["t_end"]
    host = make(fx, device_loop=False)
    dev = make(fx, device_loop=True, graph_steps=4)
    host.perform_iterations(3)
    dev.perform_iterations(3)
    t_stop = dev.time + 2e-14
    assert t_end > t_stop
    host.perform_time_evolution(t_stop)        # one step of dt = 2e-14
    dev.perform_time_evolution(t_stop)
    same_run(host, dev)
    assert dev.n_step == 4
    host.perform_time_evolution(t_stop + 5e-15)  # clamped dt < 1e-14: no step
    dev.perform_time_evolution(t_stop + 5e-15)
    same_run(host, dev)
    assert dev.n_step == 4
    assert dev.dm.dt < 1e-14
