"""Particle removal (fluidsimulation.cpp:1985-1988, 3213-3243) in the
oracle: SplitMix64 known answers (the published first outputs of
splitmix64.c for state 0), and the removal against an independent numpy
brute force: no cell above the cap, cells at or under it untouched, an
over-full cell keeps exactly its cap smallest keys, solid cells emptied,
particles outside the grid kept."""
import numpy as np
import pytest

M64 = (1 << 64) - 1


def splitmix(seed, i):
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_removal_key_known_answers(orc):
    assert orc.removal_key(0, 0) == 0xE220A8397B1DCDAF
    assert orc.removal_key(0, 1) == 0x6E789E6AA1B965F4
    assert orc.removal_key(0, 2) == 0x06C45D188009454F
    for seed, i in ((7, 0), (123456789, 99), (M64, 12345)):
        assert orc.removal_key(seed, i) == splitmix(seed, i)


def crowded_scene(seed=3, dims=(10, 8, 6), dx=0.1):
    """Cells with up to 3x the cap, a few solid cells, particles outside."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = dims
    pos = [rng.uniform(0, 1, size=(2000, 3)) * np.array(dims) * dx]
    hot = [(1, 1, 1), (4, 3, 2), (8, 6, 4)]
    for (i, j, k), m in zip(hot, (75, 31, 12)):
        pos.append((np.array([i, j, k]) + rng.uniform(0.01, 0.99, size=(m, 3))) * dx)
    pos.append(rng.uniform(-0.5, -0.1, size=(20, 3)))
    pos = np.concatenate(pos).astype(np.float32)
    solid = np.zeros((nk, nj, ni), np.uint8)
    solid[0, :, :] = 1
    solid[2, 3, 4] = 1
    return dims, dx, pos, solid


def brute_keep(pos, dims, dx, solid, cap, seed):
    ni, nj, nk = dims
    c = np.floor(pos.astype(np.float64) / dx).astype(np.int64)
    inside = ((c >= 0) & (c < np.array(dims))).all(1)
    keep = np.ones(len(pos), bool)
    flat = np.where(inside, c[:, 0] + ni * (c[:, 1] + nj * c[:, 2]), -1)
    for cell in np.unique(flat[inside]):
        idx = np.flatnonzero(flat == cell)
        k, j, i = cell // (ni * nj), (cell // ni) % nj, cell % ni
        if solid is not None and solid[k, j, i]:
            keep[idx] = False
        elif len(idx) > cap:
            keys = [splitmix(seed, int(p)) for p in idx]
            order = np.argsort(np.array(keys, dtype=np.uint64))
            keep[idx] = False
            keep[idx[order[:cap]]] = True
    return keep, flat


@pytest.mark.parametrize("cap,seed", [(20, 0), (8, 99), (1, 5)])
def test_removal_matches_brute_force(orc, cap, seed):
    dims, dx, pos, solid = crowded_scene()
    keep = orc.remove_particles(pos, dims, dx, solid, cap, seed).astype(bool)
    exp, flat = brute_keep(pos, dims, dx, solid, cap, seed)
    assert np.array_equal(keep, exp)
    counts = np.bincount(flat[keep & (flat >= 0)], minlength=int(np.prod(dims)))
    assert counts.max() <= cap
    assert keep[flat < 0].all()                       # outside the grid: kept
    in_solid = (flat >= 0) & (solid.reshape(-1)[flat.clip(0)] != 0)
    assert in_solid.any() and not keep[in_solid].any()
    assert (~keep).sum() > 0


def test_removal_without_overfull_cells_keeps_everything(orc):
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 0.8, size=(500, 3)).astype(np.float32)
    keep = orc.remove_particles(pos, (8, 8, 8), 0.1, None, 100, 0)
    assert keep.all()
