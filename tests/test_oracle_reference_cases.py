"""The reference's own CPU test cases, run against the oracle.

Each test names the reference test it re-encodes (Tests/SatinComputeRasteriserTests/…).
"""
import numpy as np
import pytest

from _pkg import orc, scr
from fixtures import jittered_cloud, lattice


def _packed_cloud(pos, col, ppb):
    return orc.Cloud.from_packed(orc.pack(pos, col, pointsPerBatch=ppb))


# NearestPointReferenceTests.swift:5-23
def test_nearest_reference_uses_reverse_z_depth():
    pos = [(0.0, 0.0, 0.25), (0.0, 0.0, 0.75)]
    col = [(1, 0, 0, 1), (0, 1, 0, 1)]
    packed = orc.pack(pos, col, pointsPerBatch=2)
    res = orc.nearest_cpu_reference(orc.Cloud.from_packed(packed), 4, 4)
    center = 1 * 4 + 2
    assert res["indices"][center] == 1
    assert packed.colors[res["indices"][center]] == 0xFF00FF00


# NearestPointReferenceTests.swift:25-43
def test_nearest_reference_keeps_lowest_index_on_tie():
    pos = [(0.0, 0.0, 0.5), (0.0, 0.0, 0.5)]
    col = [(1, 0, 0, 1), (0, 1, 0, 1)]
    packed = orc.pack(pos, col, pointsPerBatch=2)
    res = orc.nearest_cpu_reference(orc.Cloud.from_packed(packed), 4, 4)
    center = 1 * 4 + 2
    assert res["indices"][center] == 0
    assert packed.colors[0] == 0xFF0000FF


# NearestPointReferenceTests.swift:45-62
def test_nearest_reference_rejects_background_and_out_of_clip():
    pos = [(2.0, 0.0, 0.5), (0.0, 0.0, -0.5)]
    col = [(1, 0, 0, 1), (0, 1, 0, 1)]
    res = orc.nearest_cpu_reference(_packed_cloud(pos, col, 2), 4, 4)
    assert np.all(res["indices"] == 0xFFFFFFFF)
    assert np.all(res["depths"] == 0)


# LayoutTests.swift:5-26
def test_layout_strides_and_offsets():
    for dt in (scr.RASTER_BATCH_DTYPE, orc.RASTER_BATCH_DTYPE):
        assert dt.itemsize == 64
        assert dt.fields["state"][1] == 0
        assert dt.fields["minX"][1] == 4
        assert dt.fields["maxZ"][1] == 24
        assert dt.fields["numPoints"][1] == 28
        assert dt.fields["firstPoint"][1] == 32
        assert dt.fields["fileIndex"][1] == 36
    assert scr.RASTER_FILE_DTYPE.itemsize == 256
    assert scr.VISIBLE_BATCH_DTYPE.itemsize == 16


# LayoutTests.swift:33-42
def test_cube_grid_fixture_counts():
    pos, col = orc.cube_grid(4)
    packed = orc.pack(pos, col)
    assert packed.pointCount == 64
    assert packed.xyzLow.size == packed.xyzMed.size == packed.xyzHigh.size == 64
    assert packed.batchCount >= 1
    assert len(packed.files) == 1


# LayoutTests.swift:69-102
def test_pack_assigns_lod_levels_covering_all_points():
    g = np.arange(8, dtype=np.float32)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    col = np.ones((pos.shape[0], 4), np.float32)
    packed = orc.pack(pos, col, pointsPerBatch=128, lodLevels=4, coarseVoxelDivisions=2)
    counts = np.bincount(packed.levels, minlength=8)
    assert counts[:4].sum() == pos.shape[0]
    assert 0 < counts[0] < pos.shape[0]
    assert np.all(packed.colors >> 24 == 0xFF)


# LayoutTests.swift:104-134
def test_pack_preserves_color_multiset():
    pos = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]
    col = [(1, 0, 0, 1), (0, 1, 0, 1), (0, 0, 1, 1), (1, 1, 0, 1), (1, 1, 1, 1)]
    packed = orc.pack(pos, col, pointsPerBatch=2)
    assert packed.pointCount == 5 and packed.batchCount == 3
    assert sorted(packed.colors.tolist()) == sorted([0xFF0000FF, 0xFF00FF00, 0xFFFF0000, 0xFF00FFFF, 0xFFFFFFFF])


# LODBucketingTests.swift:50-59
def test_lod_cumulative_counts_round_trip():
    b = np.zeros(1, scr.RASTER_BATCH_DTYPE)[0]
    assert scr.lod_cumulative_counts(b) == [0] * 8 and b["lodCum67"] == 0
    scr.set_lod_cumulative_counts(b, [1, 5, 5, 100, 100, 100, 700, 800])
    assert scr.lod_cumulative_counts(b) == [1, 5, 5, 100, 100, 100, 700, 800]
    assert b["lodCum67"] != 0
    with pytest.raises(ValueError):
        scr.set_lod_cumulative_counts(b, [0, 0, 0, 0, 0, 0, 0, 70000])


def _lod_fixtures():
    pos, col = jittered_cloud()
    a = orc.pack(pos, col, pointsPerBatch=512, lodLevels=4, coarseVoxelDivisions=8)
    gp, gc = orc.cube_grid(24)
    b = orc.pack(gp, gc)
    return [a, b]


# LODBucketingTests.swift:61-81
def test_pack_buckets_batches_by_level_with_prefix_counts():
    for packed in _lod_fixtures():
        assert packed.batchCount > 0
        for b in packed.batches:
            f, n = int(b["firstPoint"]), int(b["numPoints"])
            sl = packed.levels[f:f + n]
            assert np.all(np.diff(sl.astype(int)) >= 0)
            recount = np.cumsum(np.bincount(sl & 7, minlength=8)).tolist()
            assert scr.lod_cumulative_counts(b) == recount
            assert recount[7] == n
            assert b["lodCum67"] != 0


# LODBucketingTests.swift:83-114
def test_pack_permutation_consistent_after_bucketing():
    pos, col = jittered_cloud(2000, 0xB0CE7)
    packed = orc.pack(pos, col, pointsPerBatch=512, lodLevels=4, coarseVoxelDivisions=8)
    src = packed.sourceIndices.astype(np.int64)
    assert len(set(src.tolist())) == len(pos)
    assert np.array_equal(packed.orderedPositions, pos[src])
    c = np.clip(col[src], 0, 1) * np.float32(255)
    expect = (c[:, 0].astype(np.uint32) | (c[:, 1].astype(np.uint32) << 8) | (c[:, 2].astype(np.uint32) << 16)
              | (c[:, 3].astype(np.uint32) << 24))
    assert np.array_equal(packed.colors, expect)


def _lmax(t):
    return min(max(int(np.floor(np.float32(t) + np.float32(0.5))), 0), 7)


# LODBucketingTests.swift:116-153
def test_survivor_prefix_is_exact_upper_bound():
    for packed in _lod_fixtures():
        for thr in (0, 0.4, 0.5, 1.0, 1.7, 3.0, 99):
            lm = _lmax(thr)
            for b in packed.batches:
                active = scr.lod_cumulative_counts(b)[lm]
                f, n = int(b["firstPoint"]), int(b["numPoints"])
                tail = packed.levels[f + active:f + n].astype(np.float32)
                assert not np.any(tail < np.float32(thr) + np.float32(0.5))
    assert _lmax(99) == 7
    for packed in _lod_fixtures():
        for b in packed.batches:
            assert scr.lod_cumulative_counts(b)[7] == b["numPoints"]


# GPUPackerParityTests.swift:55-73 — a lattice has distinct 30-bit keys, so
# the (stable) sort order equals the plain key order.
def test_lattice_has_distinct_keys_and_decodes_within_one_step():
    pos, col = lattice(16)
    packed = orc.pack(pos, col)
    keys = orc.morton_keys(pos, 30)
    assert len(np.unique(keys)) == len(keys)
    # decode at full precision reproduces every position to binary32 resolution
    # (the normalised coordinate is a float, so q carries 24 significant bits)
    b = packed.batches[0]
    mn = np.array([b["minX"], b["minY"], b["minZ"]], np.float64)
    mx = np.array([b["maxX"], b["maxY"], b["maxZ"]], np.float64)
    size = np.maximum(mx - mn, 1e-6)
    q = np.stack([((packed.xyzLow >> (10 * a)) & 1023).astype(np.int64) << 20
                  | ((packed.xyzMed >> (10 * a)) & 1023).astype(np.int64) << 10
                  | ((packed.xyzHigh >> (10 * a)) & 1023).astype(np.int64) for a in range(3)], 1)
    dec = q * (size / 2.0 ** 30) + mn
    assert np.abs(dec - packed.orderedPositions).max() <= size.max() * 2.0 ** -23
