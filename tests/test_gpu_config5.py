"""The 4M-particle galaxy (1.6M disk, 0.4M bulge, 2M halo, seed 6) run for
200 leapfrog steps at theta 0.75 and 0.5 with the learned list-length
predictor (trained on the first 50 evaluations), the rest of the steps from
captured graphs.  Checks the run physically (energy drift of the leapfrog at
dt 0.01, centre-of-mass drift) and the predictor's statistics (overflow rate,
error); list identity with the other slot policies is tested on smaller
systems (test_gpu_predict.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def galaxy4m():
    import gt_ics
    m = gt_ics.model(disk_n=1_600_000, bulge_n=400_000, halo_n=2_000_000, bulge_mass=0.333,
                     bulge_a=0.2, bulge_rmax=2.0, halo_mass=5.8, halo_rmax=30.0,
                     halo_profile="hernquist", halo_a=4.0, seed=6)
    pos, vel, mass, _ = gt_ics.galaxy(m)
    return pos, vel, mass


def energy(pos, vel, mass, pot):
    return 0.5 * np.sum(mass * np.sum(vel * vel, axis=1)) + 0.5 * np.sum(mass * pot)


@pytest.mark.parametrize("theta", [0.75, 0.5])
def test_galaxy_4m_200_steps_with_predictor(hip, galaxy4m, theta):
    pos, vel, mass = galaxy4m
    cfg = dict(theta=theta, eps=0.025, group_size=32, pack=1, predict=1)
    hip.predictor_reset(50)
    hip.state_set(pos, vel, mass)
    hip.prime(**cfg)
    _, _, _, pot0 = hip.state_get()
    e0 = energy(pos, vel, mass, pot0)
    p0 = np.sum(mass[:, None] * pos, axis=0) / mass.sum()
    before = hip.pipeline_stats()
    hip.steps(200, 0.01, use_graph=True, **cfg)
    after = hip.pipeline_stats()
    p, v, _, phi = hip.state_get()
    e1 = energy(p, v, mass, phi)
    st = hip.predictor_stats()
    print("theta", theta, "energy drift", (e1 - e0) / abs(e0), st,
          {k: after[k] - before[k] for k in after})
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert abs(e1 - e0) / abs(e0) < 2e-3
    com = np.sum(mass[:, None] * p, axis=0) / mass.sum()
    assert np.linalg.norm(com - p0) < 1e-2
    assert st["trained"] == 1 and st["predicted"] >= 100
    assert st["overflows"] <= 0.1 * st["predicted"]
    assert st["mean_abs_log_error"] < 0.2
    assert after["graph_steps"] - before["graph_steps"] >= 100


def test_what_the_trained_network_gains(hip, galaxy4m):
    """The 4M galaxy at theta 0.75 under the three slot policies, on the same
    trajectory (the states are bit-identical whichever runs): the plain
    two-pass solver (count pass, scan, fill pass), uniform slots (the last
    evaluation's longest lists + 25% for every group) and slots sized by the
    trained network (first 50 evaluations gather samples).  Reported: lists
    phase and step time, entries the slots reserve, the overflow rate.  The
    network must reserve less than uniform slots and beat the two-pass lists
    phase."""
    pos, vel, mass = galaxy4m
    base = dict(theta=0.75, eps=0.025, group_size=32, pack=1)
    res = {}
    for name, predict in (("two-pass", 2), ("uniform", 0), ("predicted", 1)):
        cfg = dict(base, predict=predict)
        if predict == 1:
            hip.predictor_reset(50)
        hip.state_set(pos, vel, mass)
        hip.prime(**cfg)
        hip.steps(55, 0.01, use_graph=True, **cfg)  # the same trajectory for all three
        s0 = hip.pipeline_stats()
        hip.steps(20, 0.01, use_graph=True, **cfg)
        s1 = hip.pipeline_stats()
        tm, inf = hip.timings(), hip.info()
        res[name] = dict(lists_ms=round(tm["lists"], 3), step_ms=round(tm["total"], 3),
                         slot_entries=s1["last_slot_entries"],
                         list_entries=inf["n_cell_list"] + inf["n_body_list"],
                         rewrites=s1["list_rewrites"] - s0["list_rewrites"],
                         graph_steps=s1["graph_steps"] - s0["graph_steps"])
    st = hip.predictor_stats()
    for name, r in res.items():
        print(f"{name:>9}: lists {r['lists_ms']:.3f} ms, step {r['step_ms']:.3f} ms, slots "
              f"{r['slot_entries'] / 1e6:.1f}M entries for {r['list_entries'] / 1e6:.1f}M "
              f"({r['slot_entries'] / r['list_entries']:.2f}x), rewrites {r['rewrites']}",
              flush=True)
    print("predictor:", st, flush=True)
    p, u, t = res["predicted"], res["uniform"], res["two-pass"]
    assert p["list_entries"] == u["list_entries"] == t["list_entries"]
    assert st["trained"] == 1 and st["predicted"] >= 20
    assert st["overflows"] <= 0.1 * st["predicted"]
    assert p["slot_entries"] < u["slot_entries"]
    assert p["lists_ms"] < t["lists_ms"]
