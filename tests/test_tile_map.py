"""The BRGEMM workgroup -> tile map (tile_of in mpsdyn_amd/csrc/zgemm_f64.hip), restated: it
must be a bijection from workgroup ids onto (row tile, column tile, batch group) for every
grid shape, whether it takes the 2D XCD blocks or the plain walk, and for the D = 512 stages
each XCD's range must be one compact block.  The device code itself is covered by every GPU
parity test: a tile the map skipped would leave its outputs unwritten."""
import pytest

K_GROUP_M = 8


def tile_of(total, bid, gm, gn):
    i = bid
    if total % 8 == 0:
        i = (i & 7) * (total >> 3) + (i >> 3)
    per_z = gm * gn
    z, r = divmod(i, per_z)
    px = py = 0
    if total % 8 == 0 and per_z % 8 == 0:
        best = 1 << 30
        for f in range(4):
            qx, qy = 8 >> f, 1 << f
            if gm % qx == 0 and gn % qy == 0 and gm // qx + gn // qy < best:
                best, px, py = gm // qx + gn // qy, qx, qy
    bm, bn, om, on = gm, gn, 0, 0
    if px:
        bm, bn = gm // px, gn // py
        blk, r = divmod(r, bm * bn)
        om, on = (blk % px) * bm, (blk // px) * bn
    g, rr = divmod(r, K_GROUP_M * bn)
    first_m = g * K_GROUP_M
    gsize = min(bm - first_m, K_GROUP_M)
    return om + first_m + rr % gsize, on + rr // gsize, z


@pytest.mark.parametrize("nz", [1, 2, 3])
def test_bijection(nz):
    for gm in range(1, 50):
        for gn in range(1, 33):
            total = gm * gn * nz
            seen = {tile_of(total, b, gm, gn) for b in range(total)}
            assert len(seen) == total
            assert all(0 <= t < gm and 0 <= u < gn and 0 <= z < nz for t, u, z in seen)


# env x site at D = 512, w = 3: one 12 x 8 block per XCD; the last stage (3 parts, 2 per
# workgroup, 2 groups): each XCD's range is two adjacent 8 x 8 blocks of one group
@pytest.mark.parametrize("gm,gn,nz,rows,cols", [(48, 16, 1, 12, 8), (32, 16, 2, 16, 8)])
def test_xcd_blocks(gm, gn, nz, rows, cols):
    total = gm * gn * nz
    for xcd in range(8):
        tiles = [tile_of(total, b, gm, gn) for b in range(xcd, total, 8)]
        assert len({t for t, _, _ in tiles}) <= rows and len({u for _, u, _ in tiles}) <= cols
