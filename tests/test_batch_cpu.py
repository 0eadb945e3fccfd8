"""Batched form: the seeded many-block generator (CPU)."""
import numpy as np
import pytest

from hidem_amd import _abi as A


@pytest.fixture(scope="module")
def small_batch():
    return A.gen_batch(n_blocks=12, min_particles=500, max_particles=4000, seed=3)


def test_block_sizes_and_ranges(small_batch):
    b = small_batch
    counts = np.array([b.blocks[k].particleCount for k in range(12)])
    starts = np.array([b.blocks[k].particleStart for k in range(12)])
    assert np.all(starts == np.concatenate([[0], np.cumsum(counts)[:-1]]))
    assert b.particles.n == counts.sum()
    assert np.all(counts >= 500 * 0.5) and np.all(counts <= 4000 * 1.6)
    for k in range(12):
        s, c = starts[k], counts[k]
        assert np.all(b.block_id[s:s + c] == k)


def test_blocks_do_not_touch(small_batch):
    b = small_batch
    x, y, z = b.particles["xCord"], b.particles["yCord"], b.particles["zCord"]
    boxes = []
    for k in range(12):
        m = b.block_id == k
        boxes.append((x[m].min(), x[m].max(), y[m].min(), y[m].max()))
    for i in range(12):
        for j in range(i + 1, 12):
            a, c = boxes[i], boxes[j]
            sep = max(c[0] - a[1], a[0] - c[1], c[2] - a[3], a[2] - c[3])
            assert sep > 1.2 + 0.5  # more than one search distance apart
    # links never cross blocks
    assert np.all(b.block_id[b.links.a] == b.block_id[b.links.b])


def test_per_block_thresholds_and_waterlines(small_batch):
    b = small_batch
    thr = np.array([b.blocks[k].threshold for k in range(12)])
    assert np.all((thr >= 2e-4) & (thr <= 5e-4))
    lb = b.block_id[b.links.a]
    ratio = b.links.maxLoad / thr[lb]
    assert np.all((ratio >= 0.9) & (ratio <= 1.1))
    for k in range(12):
        m = b.block_id == k
        z = b.particles["zCord"][m]
        frac = (b.water[k] - z.min()) / (z.max() - z.min())
        assert 0.1 < frac < 0.9


def test_batch_is_deterministic():
    b1 = A.gen_batch(n_blocks=4, min_particles=500, max_particles=2000, seed=5)
    b2 = A.gen_batch(n_blocks=4, min_particles=500, max_particles=2000, seed=5)
    assert np.array_equal(b1.particles["xCord"], b2.particles["xCord"])
    assert np.array_equal(b1.links.maxLoad, b2.links.maxLoad)


def test_full_size_layout_is_about_12m():
    A._declare_batch(A.lib())
    spec = A.HdmBatchSpec(256, 5000, 200000, 1.4707821048680188, 0.5, 0.01, 3.0, 2e-4, 5e-4, 0.1, 0.2, 0.8, 2.0, 11)
    blocks = (A.HdmBatchBlock * 256)()
    n = A.C.c_int64()
    assert A.lib().hdm_gen_batch_layout(A.C.byref(spec), blocks, A.C.byref(n)) == 0
    assert 9e6 < n.value < 16e6
