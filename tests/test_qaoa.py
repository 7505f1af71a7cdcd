"""Batched QAOA MaxCut.

CPU: the instance generator, and the oracle's QAOA (per-gate H, RZZ, RX on
the CPU oracle, then sum |a|^2 cut) against the closed-form p=1 expectation
of Wang, Hadfield, Jiang, Rieffel, PRA 97, 022304 (2018), which pins it.
GPU: the batched hipGraph run against that oracle (complex64 state: relative
tolerance 1e-5 on <C>), replay determinism, new angles without re-capture,
every tile shape giving the same answer.
"""
import math

import numpy as np
import pytest

from hisvsim_amd.qaoa import (Instance, QAOABatch, circuit, cut_values, expectation_p1, make_batch,
                              random_3regular)


def oracle_expectation(orc, inst: Instance):
    a = np.zeros(2 ** inst.n, complex)
    a[0] = 1
    orc.apply_gates(a, circuit(inst))
    return float(np.sum(np.abs(a) ** 2 * cut_values(inst.n, inst.edges))), a


def test_generator_graphs():
    rng = np.random.default_rng(3)
    for n in range(4, 23):
        e = random_3regular(n, rng)
        deg = np.zeros(n, int)
        for u, v in e:
            assert u != v
            deg[u] += 1
            deg[v] += 1
        assert len(set(e)) == len(e)
        if n % 2:
            assert sorted(deg.tolist()) == [2] + [3] * (n - 1)
        else:
            assert set(deg.tolist()) == {3}
    b = make_batch(64, seed=1)
    assert all(12 <= i.n <= 22 and 1 <= i.p <= 4 for i in b)
    assert all(0 <= x < math.pi for i in b for x in i.gammas + i.betas)


@pytest.mark.parametrize("n", [4, 7, 10, 12])
def test_oracle_matches_closed_form_p1(orc, n):
    rng = np.random.default_rng(n)
    for _ in range(3):
        edges = random_3regular(n, rng)
        g, b = rng.uniform(0, math.pi, 2)
        inst = Instance(n, edges, [float(g)], [float(b)])
        got, _ = oracle_expectation(orc, inst)
        assert abs(got - expectation_p1(n, edges, g, b)) <= 1e-12


def test_closed_form_with_triangles(orc):
    # K4 is 3-regular and every edge lies on two triangles
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    inst = Instance(4, edges, [1.1], [0.3])
    got, _ = oracle_expectation(orc, inst)
    assert abs(got - expectation_p1(4, edges, 1.1, 0.3)) <= 1e-12


@pytest.mark.gpu
def test_batch_matches_oracle(orc):
    rng = np.random.default_rng(7)
    insts = make_batch(48, seed=7, n_range=(4, 15), p_range=(0, 4))
    with QAOABatch(insts) as b:
        got = b.run()
        for i, inst in enumerate(insts):
            want, amps = oracle_expectation(orc, inst)
            assert abs(got[i] - want) <= 1e-5 * max(1.0, abs(want)), (i, inst.n, inst.p)
        # the state itself, for a few instances
        for i in (0, 5, 17):
            _, amps = oracle_expectation(orc, insts[i])
            assert np.abs(b.state(i) - amps).max() <= 2e-6


@pytest.mark.gpu
def test_batch_large_instances(orc):
    insts = make_batch(6, seed=11, n_range=(20, 22), p_range=(1, 2))
    with QAOABatch(insts) as b:
        got = b.run()
    for i in (0, 3):
        want, _ = oracle_expectation(orc, insts[i])
        assert abs(got[i] - want) <= 1e-5 * abs(want)


@pytest.mark.gpu
def test_graph_replay_and_new_angles(orc):
    insts = make_batch(32, seed=2, n_range=(10, 16), p_range=(1, 3))
    with QAOABatch(insts) as b:
        r1 = b.run()
        r2 = b.run()
        assert np.array_equal(r1, r2)  # deterministic reductions, same graph replayed
        new = [Instance(i.n, i.edges, [g * 0.5 for g in i.gammas], i.betas) for i in insts]
        b.set_angles([i.gammas for i in new], [i.betas for i in new])
        r3 = b.run()
        for i in (0, 9):
            want, _ = oracle_expectation(orc, new[i])
            assert abs(r3[i] - want) <= 1e-5 * abs(want)


@pytest.mark.gpu
def test_tile_shapes_agree():
    insts = make_batch(24, seed=4, n_range=(12, 20), p_range=(1, 2))
    with QAOABatch(insts) as b:
        ref = b.run()
        for cb, t, g in [(11, 128, 2), (12, 512, 8), (13, 256, 4)]:
            b.tune("chunk_bits", cb)
            b.tune("threads", t)
            b.tune("grid_per_cu", g)
            assert np.abs(b.run() - ref).max() <= 1e-5 * np.abs(ref).max()
