"""Slot pool (SlotPoolTests.swift) and GPU packer (GPUPackerParityTests.swift)
through the C ABI.  The GPU pack must equal the oracle's CPU pack bit for bit:
same stable Morton order, LOD levels, bucketing, batch table and quantised words."""
import numpy as np
import pytest

from _pkg import orc, scr
from fixtures import jittered_cloud, lattice

pytestmark = pytest.mark.gpu


def _batches(n, num=4, first=None):
    b = np.zeros(n, scr.RASTER_BATCH_DTYPE)
    b["maxX"] = b["maxY"] = b["maxZ"] = 1.0
    b["numPoints"] = num
    b["state"] = 1
    b["firstPoint"] = first if first is not None else 0
    return b


# SlotPoolTests.swift:23-36
def test_streaming_init_allocates_empty_pool(ctx):
    c = scr.ComputeRasteriserPointCloud(ctx, 8, 16)
    assert c.batchCount == 8 and c.residentBatchCount == 0 and c.pointCount == 0 and c.freeSlotCount == 8
    assert c.readPool(scr.POOL_XYZ_LOW).size == 8 * 16
    assert np.all(c.readPool(scr.POOL_BATCHES)["state"] == 0)


# SlotPoolTests.swift:38-86
def test_add_batches_assigns_slots_lifo_and_rebases(ctx):
    c = scr.ComputeRasteriserPointCloud(ctx, 4, 8)
    pos = np.arange(8, dtype=np.uint32) + 100
    col = np.arange(8, dtype=np.uint32) + 200
    lev = np.arange(8, dtype=np.uint8)
    b = _batches(2)
    b["firstPoint"] = [0, 4]
    slots = c.addBatches(pos, pos, pos, col, lev, b)
    assert slots == [0, 1]
    assert c.residentBatchCount == 2 and c.pointCount == 8 and c.freeSlotCount == 2
    m = c.readPool(scr.POOL_BATCHES)
    assert list(m["state"]) == [1, 1, 0, 0] and list(m["firstPoint"][:2]) == [0, 8]
    assert list(m["numPoints"][:2]) == [4, 4]
    low = c.readPool(scr.POOL_XYZ_LOW)
    assert list(low[0:4]) == list(pos[0:4]) and list(low[8:12]) == list(pos[4:8])
    lv = c.readPool(scr.POOL_LEVELS)
    assert list(lv[0:4]) == list(lev[0:4]) and list(lv[8:12]) == list(lev[4:8])
    assert list(c.readPool(scr.POOL_COLORS)[8:12]) == list(col[4:8])


# SlotPoolTests.swift:88-121
def test_remove_batches_frees_slots_and_reuses_lifo(ctx):
    c = scr.ComputeRasteriserPointCloud(ctx, 4, 8)
    z = np.zeros(8, np.uint32)
    b = _batches(2)
    b["firstPoint"] = [0, 4]
    slots = c.addBatches(z, z, z, z, np.zeros(8, np.uint8), b)
    c.removeBatches([slots[0]])
    assert c.residentBatchCount == 1 and c.pointCount == 4 and c.freeSlotCount == 3
    m = c.readPool(scr.POOL_BATCHES)
    assert m["state"][slots[0]] == 0 and m["numPoints"][slots[0]] == 0 and m["state"][slots[1]] == 1
    assert c.addBatches(z, z, z, z, np.zeros(8, np.uint8), _batches(1)) == [slots[0]]


# SlotPoolTests.swift:123-166
def test_deferred_commit_flushes_only_dirty_slots(ctx):
    c = scr.ComputeRasteriserPointCloud(ctx, 8, 8)
    z = np.zeros(8, np.uint32)
    slots = c.addBatches(z, z, z, z, np.zeros(8, np.uint8), _batches(8), commit=False)
    assert slots == list(range(8))
    assert np.all(c.readPool(scr.POOL_BATCHES)["state"] == 0)
    c.commitBatchUpdates()
    assert np.all(c.readPool(scr.POOL_BATCHES)["state"] == 1)
    c.removeBatches([6], commit=False)
    c.commitBatchUpdates()
    m = c.readPool(scr.POOL_BATCHES)
    assert m["state"][6] == 0 and m["numPoints"][6] == 0
    for s in (0, 5, 7):
        assert m["state"][s] == 1 and m["numPoints"][s] == 4 and m["firstPoint"][s] == s * 8
    c.commitBatchUpdates()
    assert c.readPool(scr.POOL_BATCHES)["state"][7] == 1


# SlotPoolTests.swift:168-176
def test_packed_init_fits_capacity_to_batch_count(ctx):
    pos, col = orc.cube_grid(8)
    packed = orc.pack(pos, col)
    c = scr.ComputeRasteriserPointCloud.from_packed(ctx, packed)
    assert c.capacity[0] == packed.batchCount and c.residentBatchCount == packed.batchCount
    assert c.pointCount == packed.pointCount


def test_pool_errors_are_status_codes(ctx):
    c = scr.ComputeRasteriserPointCloud(ctx, 2, 8)
    z = np.zeros(16, np.uint32)
    lv = np.zeros(16, np.uint8)
    with pytest.raises(scr.ScrError) as e:
        c.addBatches(z, z, z, z, lv, _batches(3))
    assert e.value.code == -2
    with pytest.raises(scr.ScrError) as e:
        c.addBatches(z, z, z, z, lv, _batches(1, num=9))
    assert e.value.code == -3
    b = _batches(1)
    b["firstPoint"] = 14
    with pytest.raises(scr.ScrError) as e:
        c.addBatches(z, z, z, z, lv, b)
    assert e.value.code == -7
    with pytest.raises(scr.ScrError) as e:
        c.removeBatches([5])
    assert e.value.code == -9
    assert c.residentBatchCount == 0 and c.freeSlotCount == 2  # failed calls changed nothing
    with pytest.raises(scr.ScrError) as e:
        scr.ComputeRasteriserPointCloud(ctx, 2, 70000)
    assert e.value.code == -4


# ---------------------------------------------------------------- GPU pack parity
def _assert_pack_equal(cloud, ref, first_slot=0):
    n = ref.pointCount
    ppb = cloud.capacity[1]
    base = first_slot * ppb
    nb = ref.batchCount
    gb = cloud.readPool(scr.POOL_BATCHES)[first_slot:first_slot + nb].copy()
    rb = ref.batches.copy()
    rb["firstPoint"] = rb["firstPoint"] + base
    assert gb.tobytes() == rb.tobytes(), (gb[:2], rb[:2])
    for which, arr in ((scr.POOL_XYZ_LOW, ref.xyzLow), (scr.POOL_XYZ_MED, ref.xyzMed),
                       (scr.POOL_XYZ_HIGH, ref.xyzHigh), (scr.POOL_COLORS, ref.colors),
                       (scr.POOL_LEVELS, ref.levels)):
        got = cloud.readPool(which)[base:base + n]
        bad = np.flatnonzero(got != arr)
        assert bad.size == 0, (which, bad[:5], got[bad[:5]], arr[bad[:5]])


def _pack_cases():
    cases = []
    p, c = lattice(16)
    cases.append(("lattice16", p, c, 10240, 4, 64))
    p, c = lattice(40)
    cases.append(("lattice40", p, c, 10240, 4, 64))
    p, c = jittered_cloud(20000, 0x5EED)
    p[100:300] = p[7]                        # equal keys: stability
    cases.append(("jitter_ties", p, c, 777, 4, 8))
    g = np.arange(8, dtype=np.float32)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    cases.append(("grid8_coarse2", p, np.ones((512, 4), np.float32), 128, 4, 2))
    p, c = jittered_cloud(5000, 0xABC)
    cases.append(("lod1", p, c, 1000, 1, 64))
    cases.append(("lod8_coarse2", p, c, 1000, 8, 2))
    return cases


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("case", _pack_cases(), ids=lambda c: c[0])
def test_gpu_pack_equals_cpu_pack(ctx, case, bits):
    name, pos, col, ppb, lod, coarse = case
    ref = orc.pack(pos, col, pointsPerBatch=ppb, lodLevels=lod, coarseVoxelDivisions=coarse, keyBits=bits)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, ppb)
    packer = scr.GPUPacker(ctx, lodLevels=lod, coarseVoxelDivisions=coarse, keyBits=bits)
    src = packer.pack(cloud, pos, col, sourceIndices=True)
    assert cloud.residentBatchCount == ref.batchCount and cloud.pointCount == ref.pointCount
    assert np.array_equal(src, ref.sourceIndices)
    _assert_pack_equal(cloud, ref)


def test_ply_input_packs_like_the_cpu_pack(ctx):
    # PLYPointCloudLoader hands its arrays to the pack: the reference's own
    # case (LayoutTests.swift:44-67) and a binary file of a jittered cloud
    ply = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
           "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n"
           "0 0 0 255 0 0\n1 0 0 0 255 0\n0 1 0 0 0 255").encode()
    p, c = jittered_cloud(6000, 0x91F)
    rec = np.zeros(p.shape[0], [("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("r", "u1"), ("g", "u1"), ("b", "u1")])
    rec["x"], rec["y"], rec["z"] = p[:, 0], p[:, 1], p[:, 2]
    for k, ch in enumerate("rgb"):
        rec[ch] = np.rint(np.clip(c[:, k], 0, 1) * 255).astype(np.uint8)
    head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {p.shape[0]}\nproperty float x\n"
            "property float y\nproperty float z\nproperty uchar r\nproperty uchar g\nproperty uchar b\nend_header\n")
    binary = head.encode() + rec.tobytes()
    for data, ppb in ((ply, 2), (binary, 777)):
        pos, col = scr.PLYPointCloudLoader.parse(data)
        ref = orc.pack(pos[:, :3], col, pointsPerBatch=ppb)
        cloud = scr.ComputeRasteriserPointCloud(ctx, 1, ppb)
        n = scr.PLYPointCloudLoader.packInto(scr.GPUPacker(ctx), cloud, data)
        assert n == ref.pointCount and cloud.residentBatchCount == ref.batchCount
        _assert_pack_equal(cloud, ref)
        if data is ply:
            rgb = set(int(v) & 0x00FFFFFF for v in cloud.readPool(scr.POOL_COLORS)[:3])
            assert rgb == {0x000000FF, 0x0000FF00, 0x00FF0000}


def test_gpu_pack_generated_scene_both_key_widths(ctx):
    pos, col = scr.gen_scene(ctx, 300_000, on_device=False, seed=0x5C0002)
    for bits in (30, 63):
        ref = orc.pack(pos[:, :3], col, pointsPerBatch=10240, keyBits=bits)
        cloud = scr.ComputeRasteriserPointCloud(ctx, 1, 10240)
        src = scr.GPUPacker(ctx, keyBits=bits).pack(cloud, pos, col, sourceIndices=True)
        assert np.array_equal(src, ref.sourceIndices)
        _assert_pack_equal(cloud, ref)


# GPUPackerParityTests.swift:262-323
def test_chunk_pack_slot_targeted_non_destructive(ctx):
    pa, ca = lattice(40)
    pb, cb = lattice(32)
    ra = orc.pack(pa, ca)
    rb = orc.pack(pb, cb)
    cloud = scr.ComputeRasteriserPointCloud(ctx, ra.batchCount + rb.batchCount + 3, 10240)
    packer = scr.GPUPacker(ctx)
    assert packer.addRawChunksGPU(cloud, [(pa, ca)]) == [0]
    assert cloud.residentBatchCount == ra.batchCount and cloud.pointCount == ra.pointCount
    _assert_pack_equal(cloud, ra, 0)
    assert packer.addRawChunksGPU(cloud, [(pb, cb)]) == [ra.batchCount]
    assert cloud.residentBatchCount == ra.batchCount + rb.batchCount
    _assert_pack_equal(cloud, rb, ra.batchCount)
    _assert_pack_equal(cloud, ra, 0)
    # two chunks in one call (GPUPackerParityTests.swift:325-356 shape)
    cloud2 = scr.ComputeRasteriserPointCloud(ctx, ra.batchCount + rb.batchCount + 1, 10240)
    assert packer.addRawChunksGPU(cloud2, [(pa, ca), (pb, cb)]) == [0, ra.batchCount]
    _assert_pack_equal(cloud2, ra, 0)
    _assert_pack_equal(cloud2, rb, ra.batchCount)
    # no contiguous run left: the chunk is skipped
    assert packer.addRawChunksGPU(cloud2, [(pa, ca)]) == []


def test_gen_scene_is_deterministic_and_host_equals_device(ctx):
    a_pos, a_col = scr.gen_scene(ctx, 100_000, on_device=False, seed=7)
    b_pos, b_col = scr.gen_scene(ctx, 100_000, on_device=False, seed=7)
    assert np.array_equal(a_pos, b_pos) and np.array_equal(a_col, b_col)
    dp, dc = scr.gen_scene(ctx, 100_000, on_device=True, seed=7)
    assert np.array_equal(dp.to_array(np.float32, 400_000).reshape(-1, 4), a_pos)
    # RGB8-quantised colours survive the packer's truncation
    v = a_col[:, :3] * np.float32(255)
    assert np.allclose(v - np.floor(v), 0.5, atol=1e-3)
    # 70 % terrain / 30 % box faces, inside the square
    assert np.all(np.abs(a_pos[:70_000, :2]) <= 1000.0)


@pytest.mark.parametrize("levels,coarse", [(8, 64), (6, 64), (6, 4)])
def test_gpu_pack_lod_gate(ctx, levels, coarse):
    """Above GPUPacker's grid cap (GPUPacker.swift:35, 86-99) the GPU pack
    leaves every level at the finest instead of failing; below it, LOD runs."""
    pos, col = jittered_cloud(6000, 0x6A7E)
    ref = orc.pack(pos, col, pointsPerBatch=1000, lodLevels=levels, coarseVoxelDivisions=coarse, gpuRules=True)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, 1000)
    src = scr.GPUPacker(ctx, lodLevels=levels, coarseVoxelDivisions=coarse, keyBits=30).pack(cloud, pos, col,
                                                                                         sourceIndices=True)
    assert np.array_equal(src, ref.sourceIndices)
    _assert_pack_equal(cloud, ref)
    lv = cloud.readPool(scr.POOL_LEVELS)[:6000]
    if orc.gpu_lod_enabled(levels, coarse):
        assert len(set(lv.tolist())) > 2
    else:
        assert np.all(lv == levels - 1)
