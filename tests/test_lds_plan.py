"""The persistent engine's LDS plan: which nodes and primitive records of each scene a workgroup keeps
in LDS (32-byte nodes, 16-byte primitive records; `plan_lds` in wf_persistent.hip).

A scene whose whole tree fits, smallest need first, is kept whole (then its primitive records if
they fit too) as long as 4 KiB stay for each other scene not yet whole; the others share what is
left evenly and keep the breadth-first top of their trees. The exact plans below are the batch
benchmark's scenes; the properties hold for any scene set.
"""
import numpy as np

NODE, PRIM, MIN_TOP = 32, 16, 4096
CORNELL, SPHERE_GRID, HEIGHTFIELD = (27, 13), (515, 1026), (1049265, 349525)


def plan(wf, scenes, budget):
    nn, npr = zip(*scenes)
    return wf.test_lds_plan(list(nn), list(npr), budget)


def test_single_scene_whole_or_top(wf):
    nc, pc = plan(wf, [SPHERE_GRID], 64 * 1024)
    assert (nc[0], pc[0]) == SPHERE_GRID
    nc, pc = plan(wf, [SPHERE_GRID], 20 * 1024)      # nodes fit, primitive records do not
    assert (nc[0], pc[0]) == (515, 0)
    nc, pc = plan(wf, [HEIGHTFIELD], 32 * 1024)
    assert (nc[0], pc[0]) == (1024, 0)


def test_batch_scenes_at_32_and_40_kib(wf):
    scenes = [HEIGHTFIELD, SPHERE_GRID, CORNELL]
    nc, pc = plan(wf, scenes, 40 * 1024)
    # cornell whole; sphere_grid whole with its records (4 KiB kept for heightfield); the rest
    # is heightfield's top
    left = 40 * 1024 - (27 * NODE + 13 * PRIM) - (515 * NODE + 1026 * PRIM)
    assert list(nc) == [left // NODE, 515, 27] and list(pc) == [0, 1026, 13]
    nc, pc = plan(wf, scenes, 32 * 1024)
    left = 32 * 1024 - (27 * NODE + 13 * PRIM) - 515 * NODE
    assert list(nc) == [left // NODE, 515, 27] and list(pc) == [0, 0, 13]


def test_plan_properties_on_random_scene_sets(wf):
    rs = np.random.default_rng(8)
    for _ in range(500):
        n = int(rs.integers(1, 9))
        nn = rs.integers(0, 1 << int(rs.integers(1, 21)), n)
        npr = rs.integers(0, np.maximum(nn, 1) * 4)
        budget = int(rs.integers(8, 160)) * 1024
        nc, pc = wf.test_lds_plan(nn, npr, budget)
        assert (nc * NODE + pc * PRIM).sum() <= budget
        assert ((nc >= 0) & (nc <= nn)).all()
        assert ((pc == 0) | ((pc == npr) & (nc == nn))).all()          # records only with a whole tree
        for k in range(n):
            if nc[k] < nn[k]:                                           # a partial tree keeps a top
                assert nc[k] >= min(nn[k], MIN_TOP // NODE) or budget < MIN_TOP * n
