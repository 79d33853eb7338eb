"""Launch order of the strided-GEMM launch's tasks (gecco_hip_task_order):
host logic, no GPU.  Workgroup w of the launch runs on XCD w mod 8; the
order keeps the tiles of one patch (which share operand panels) on one XCD
and spreads the patches over the XCDs by cost."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'gecco-public_amd'))

NX = 8


@pytest.fixture
def gecco_hip():
    # imported inside the test, as in the other test modules: loading the
    # library at collection time, before torch has set up the GPU, leaves the
    # library's own HIP calls without a device in the GPU runs
    import gecco_hip
    return gecco_hip


def test_hand_worked_order(gecco_hip):
    # patches by longest task: 0 (11), 1 (9), 2 (7), 3 (6), 4 (3), 5 (2) go to
    # XCDs 0..5; XCDs 6 and 7 have no patch, so positions 6 and 7 take the
    # list with most tasks left (XCD 3, then XCD 0), and so on
    perm = gecco_hip.task_order([0, 0, 1, 1, 2, 3, 3, 3, 4, 5], [10, 9, 8, 7, 6, 5, 4, 3, 2, 1])
    assert perm.tolist() == [0, 2, 4, 5, 8, 9, 6, 1, 3, 7]


def test_least_loaded_xcd_takes_the_next_patch(gecco_hip):
    # patches 0..6: three long tasks each; patch 7: one task of 50; patch 8:
    # two tasks of 10 and 9.  Patch 8 goes to XCD 7, the least loaded, so all
    # eight lists hold three tasks and position 8k + x runs task k of XCD x
    patch = [i for i in range(7) for _ in range(3)] + [7, 8, 8]
    cost = [1000 - 3 * i - j for i in range(7) for j in range(3)] + [50, 10, 9]
    perm = gecco_hip.task_order(patch, cost)
    lists = [[3 * x, 3 * x + 1, 3 * x + 2] for x in range(7)] + [[21, 22, 23]]
    assert perm.tolist() == [lists[x][k] for k in range(3) for x in range(NX)]


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_random_inputs_give_a_permutation(gecco_hip, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 5000))
    patch = rng.integers(0, max(1, n // 6), n)
    cost = np.sort(rng.integers(1, 4096, n))[::-1]
    perm = gecco_hip.task_order(patch, cost)
    assert sorted(perm.tolist()) == list(range(n))


def test_patches_stay_on_one_xcd_when_the_lists_are_even(gecco_hip):
    # 8 x 5 patches of 4 equal tasks: every XCD gets 5 patches (20 tasks), no
    # list runs out early, and each patch's tasks are consecutive on its XCD
    npatch, per = 40, 4
    patch = np.repeat(np.arange(npatch), per)
    cost = np.full(npatch * per, 256)
    perm = gecco_hip.task_order(patch, cost)
    xcd_of = {}
    seq = {x: [] for x in range(NX)}
    for w, t in enumerate(perm):
        x = w % NX
        xcd_of.setdefault(patch[t], set()).add(x)
        seq[x].append(patch[t])
    assert all(len(v) == 1 for v in xcd_of.values())
    for x in range(NX):
        assert len(seq[x]) == npatch * per // NX
        runs = [seq[x][i] for i in range(len(seq[x])) if i == 0 or seq[x][i] != seq[x][i - 1]]
        assert len(runs) == len(set(runs))          # a patch is one run on its XCD


def test_xcd_loads_balanced_longest_first(gecco_hip):
    # uneven patches: up to the first list that runs out, the XCD loads
    # differ by at most twice the largest patch
    rng = np.random.default_rng(7)
    sizes = rng.integers(1, 17, 64)
    patch = np.repeat(np.arange(64), sizes)
    cost = np.sort(rng.integers(32, 2048, len(patch)))[::-1]
    order = np.argsort(patch, kind='stable')
    patch, cost = patch[order], cost[order]
    idx = np.argsort(-cost, kind='stable')                # tasks in cost order
    patch, cost = patch[idx], cost[idx]
    perm = gecco_hip.task_order(patch, cost)
    n = len(perm)
    per_xcd = n // NX
    load = np.zeros(NX)
    for w in range(per_xcd * NX):                          # before any list runs out
        t = perm[w]
        load[w % NX] += cost[t] + 1
    ptot = {p: cost[patch == p].sum() + (patch == p).sum() for p in np.unique(patch)}
    assert load.max() - load.min() <= 2 * max(ptot.values())
    assert len(set(perm.tolist())) == n


def test_bad_arguments_raise(gecco_hip):
    with pytest.raises(gecco_hip.GeccoHipError):
        gecco_hip._check(gecco_hip._lib.gecco_hip_task_order(-1, None, None, None))
    with pytest.raises(ValueError):
        gecco_hip.task_order([0, 1], [1])
