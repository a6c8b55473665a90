"""Whole-episode parity at the benchmark sizes, with the benchmark's kernels.

* 512 x 512 fold: the full 100-substep episode under graph replay with the
  incremental sort and the tile query fused with the cloth step, once per
  refit block the autotuner can pick; at the end the state, the contacts, the
  tree and the last substep's candidate set (audited from the fused kernel)
  equal the oracle's, and no replay kept a previous topology.
* The batched scene: all 256 instances (3.0 M vertices) for its 50 substeps;
  8 instances spanning the side range equal their own plain scenes in the
  oracle (state, contacts, last substep's candidates).
* 1024 x 1024 + 4 level-6 icospheres: the LBVH candidate set at substep 0
  and at substep 4 equals the oracle's.
* The fold past the meeting of its layers (substeps 300-320, with contacts):
  the oracle restarted from the GPU state follows it bit for bit.
(Simulator::update substep order, include/sim/simulator.hpp:3512-3815.)
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from paritylib import heavy_tile_leaves, assert_same_candidates
from ysim_amd import scenes
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu


def assert_contacts_equal(cg, co, what):
    assert len(cg) == len(co), f"{what}: {len(cg)} contacts vs {len(co)}"
    for f in ("point", "triangle", "query_obj", "target_obj", "normal_distance"):
        assert np.array_equal(cg[f], co[f]), f"{what}: contact field {f}"


@pytest.fixture(scope="module")
def c2_oracle_episode(oracle_threads):
    """The oracle's 100-substep fold episode: end state, contacts, tree and the
    candidates of substep 99."""
    sc = scenes.config_c2()
    os_ = OracleScene(sc)
    steps = sc.info["substeps"]
    os_.step(steps - 1)
    pairs = os_.pairs()
    os_.step(1)
    off, con = os_.contacts()
    nodes, parent, keys, idx = os_.tree(0)
    return dict(state=os_.state(), off=off, con=con, nodes=nodes, parent=parent, keys=keys, idx=idx,
                pairs=pairs, steps=steps)


@pytest.mark.parametrize("rb", [256, 512, 1024])
def test_c2_episode_bench_tuning(ctx, c2_oracle_episode, rb):
    ref = c2_oracle_episode
    sc = scenes.config_c2()
    hs = HipScene(ctx, sc)
    hs.set_tuning(sort_digit_bits=8, sort_items=16, sort_incremental=1, sort_hybrid=0, query_tile=1,
                  refit_block=rb, build_mode=0)
    hs.set_audit(1 << 21)
    hs.step(ref["steps"], graph=True)
    deferred = hs.deferred_builds()
    print(f"refit block {rb}: deferred builds over the episode = {deferred}")
    assert deferred == 0
    for a, b in zip(hs.state(), ref["state"]):
        assert np.array_equal(a, b)
    off, con = hs.contacts()
    assert np.array_equal(off, ref["off"])
    assert_contacts_equal(con, ref["con"], f"fold rb={rb}")
    nodes, parent, keys, idx = hs.tree(0, keys=True)
    assert np.array_equal(keys, ref["keys"]) and np.array_equal(idx, ref["idx"])
    assert np.array_equal(parent, ref["parent"])
    for f in ("min", "max", "child_a", "child_b"):
        assert np.array_equal(nodes[f], ref["nodes"][f]), f
    got, _ = hs.audit()
    assert_same_candidates(got, ref["pairs"], f"fold substep {ref['steps'] - 1}, rb={rb}")
    assert hs.flags() == 0


def test_c2_contact_active_from_gpu_state(ctx, oracle_threads):
    """The fold after its layers meet: the GPU runs substeps 0-300 under the
    bench tuning, the oracle restarts from that state, and over substeps
    300-320 (contacts grow from ~70 to ~1,800 per substep, no slot overflow)
    both agree bit for bit every 5 substeps: state, contacts and, on the
    last substep, the audited candidate set of the fused kernel."""
    sc = scenes.config_c2()
    hs = HipScene(ctx, sc)
    hs.set_tuning(sort_digit_bits=8, sort_items=16, sort_incremental=1, sort_hybrid=0, query_tile=1,
                  refit_block=1024, build_mode=0)
    hs.step(300, graph=True)
    os_ = OracleScene(sc)
    os_.set_state(*hs.state())
    counts = []
    for k in range(4):
        if k == 3:
            hs.step(4, graph=True)
            os_.step(4)
            want = os_.pairs()
            # the audited substep has leaves the tile walk tests with the
            # whole wave (more than 16 row x column crossings of a tile)
            xo, vo, _ = os_.state()
            heavy = heavy_tile_leaves(xo, vo, sc.meshes[0].facets, 512, sc.params["margin"], sc.params["subh"])
            print(f"heavy (tile, leaf) pairs at substep 319: {heavy}", flush=True)
            assert heavy > 0
            hs.set_audit(1 << 21)
            hs.step(1)
            os_.step(1)
        else:
            hs.step(5, graph=True)
            os_.step(5)
        for a, b in zip(hs.state(), os_.state()):
            assert np.array_equal(a, b), f"substep {305 + 5 * k}"
        off, con = hs.contacts()
        ooff, ocon = os_.contacts()
        assert np.array_equal(off, ooff)
        assert_contacts_equal(con, ocon, f"contact-active fold, substep {305 + 5 * k}")
        counts.append(len(con))
    got, redo = hs.audit()
    assert_same_candidates(got, want, "contact-active fold, substep 319")
    print(f"contacts at substeps 305..320: {counts}")
    assert counts[0] > 0 and counts[-1] > counts[0]
    assert hs.flags() == 0


def instance_rows(rows, i):
    """Candidate rows of batched instance i (meshes 2i, 2i+1) in the instance's
    own mesh numbering."""
    r = rows[(rows[:, 0] == 2 * i) | (rows[:, 0] == 2 * i + 1)].copy()
    r[:, 0] -= 2 * i
    r[:, 2] -= 2 * i
    return r


def test_c3_full_batch_episode(ctx, oracle_threads):
    sc = scenes.config_c3()
    nI = len(sc.meshes) // 2
    steps = sc.info["substeps"]
    hs = HipScene(ctx, sc)
    hs.set_audit(1 << 24)
    hs.step(steps, graph=True)
    assert hs.flags() == 0
    x, v, xp = hs.state()
    off, con = hs.contacts()
    got, _ = hs.audit()
    sides = np.array([int(sc.meshes[2 * i].cloth["particle_num_1d"]) for i in range(nI)])
    order = np.argsort(sides, kind="stable")
    picks = [int(order[int(round(q * (nI - 1)))]) for q in np.linspace(0.0, 1.0, 8)]
    assert sides[picks[0]] <= 20 and sides[picks[-1]] >= 200
    vo, _ = sc.offsets()
    for i in picks:
        s0, s1 = int(vo[2 * i]), int(vo[2 * i + 1])
        o = OracleScene(scenes.instance_scene(sc, i))
        o.step(steps - 1)
        want = o.pairs()
        o.step(1)
        xo, vo_, po = o.state()
        n = s1 - s0
        what = f"instance {i} (side {sides[i]})"
        assert np.array_equal(x[s0:s1], xo[:n]), what
        assert np.array_equal(v[s0:s1], vo_[:n]), what
        assert np.array_equal(xp[s0:s1], po[:n]), what
        c = con[off[s0]:off[s1]].copy()
        c["query_obj"] -= 2 * i
        c["target_obj"] -= 2 * i
        assert_contacts_equal(c, o.contacts()[1], what)
        assert_same_candidates(instance_rows(got, i), want, what)


def test_c5_full_candidates(ctx, oracle_threads):
    sc = scenes.config_c5()
    hs, os_ = HipScene(ctx, sc), OracleScene(sc)
    hs.set_audit(1 << 23)
    for k in (1, 4):
        if k > 1:
            os_.step(k - 1)
        want = os_.pairs()
        os_.step(1)
        hs.step(k, graph=True)
        got, _ = hs.audit()
        assert len(want) > 2_000_000
        assert_same_candidates(got, want, f"c5 round of {k}")
    for a, b in zip(hs.state(), os_.state()):
        assert np.array_equal(a, b)
    assert hs.flags() == 0
