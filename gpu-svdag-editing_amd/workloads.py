"""Seeded synthetic workloads.

slab_hash_workload: the slab-hash microbenchmark input -- DAG interior nodes
of 2..9 words (child mask uniform over the 255 non-empty masks, child words
uniform u32), grouped by size, plus a find set with a given hit rate (hits
drawn uniformly from the inserted items, misses fresh random items with the
same size distribution).  Item seed 3, hit seed 4, miss seed 5.
"""
import numpy as np

POPC8 = np.array([bin(i).count("1") for i in range(256)], dtype=np.int64)


def _items_of_size(rng, s, n):
    masks_with = np.nonzero(POPC8 == s - 1)[0]
    out = np.empty((n, s), dtype=np.uint32)
    out[:, 0] = masks_with[rng.integers(0, len(masks_with), n)]
    if s > 1:
        out[:, 1:] = rng.integers(0, 1 << 32, size=(n, s - 1), dtype=np.uint64).astype(np.uint32)
    return out


def slab_hash_items(n_items, seed=3):
    """Returns {size: uint32[count, size]} for n_items DAG interior nodes."""
    rng = np.random.default_rng(seed)
    masks = rng.integers(1, 256, n_items)
    sizes = 1 + POPC8[masks]
    out = {}
    for s in range(2, 10):
        sel = masks[sizes == s]
        n = len(sel)
        if n == 0:
            continue
        arr = np.empty((n, s), dtype=np.uint32)
        arr[:, 0] = sel
        arr[:, 1:] = rng.integers(0, 1 << 32, size=(n, s - 1), dtype=np.uint64).astype(np.uint32)
        out[s] = arr
    return out


def slab_hash_finds(items, n_finds, hit_rate=0.5, hit_seed=4, miss_seed=5):
    """Find set with the inserted size distribution: {size: uint32[count, size]}."""
    total = sum(len(a) for a in items.values())
    rh = np.random.default_rng(hit_seed)
    rm = np.random.default_rng(miss_seed)
    out = {}
    for s, arr in items.items():
        n_s = int(round(n_finds * len(arr) / total))
        hits = int(round(n_s * hit_rate))
        h = arr[rh.integers(0, len(arr), hits)]
        m = _items_of_size(rm, s, n_s - hits)
        both = np.concatenate([h, m]) if len(m) else h
        perm = rh.permutation(len(both))
        out[s] = np.ascontiguousarray(both[perm])
    return out


def line_slots(s):
    """Items of s words per 128-byte table line: the largest k with
    k*s + ceil(k/4) <= 31 (tags, items and the chain link share the line)."""
    k = 1
    while (k + 1) * s + (k + 4) // 4 <= 31:
        k += 1
    return k


def buckets_for(items, fill=0.75):
    """Bucket count per size: the items fill `fill` of the head lines' slots,
    ceil(n_s / (line_slots(s) * fill))."""
    return {s: max(1, int(np.ceil(len(a) / (line_slots(s) * fill)))) for s, a in items.items()}


# ---------------------------------------------------------------- editing workloads
SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, BOX_REMOVE, PAINT, COPY = range(6)


def terrain_params(levels):
    """fBm heightfield terrain: 5 octaves at depth 10 (base 512, amplitude 200,
    frequency 1/256); at depth >= 15 the same shape scaled to the grid (base
    2^(D-1), amplitude 6000 * 2^(D-15), 8 octaves, 4 periods across)."""
    side = 1 << levels
    if levels < 15:
        scale = side / 1024.0
        return dict(seed=0x5EED, base=512.0 * scale, amplitude=200.0 * scale, frequency=1.0 / (256.0 * scale),
                    octaves=5, material=1)
    return dict(seed=0x5EED, base=side / 2.0, amplitude=6000.0 * side / 32768.0, frequency=4.0 / side,
                octaves=8, material=1)


def config1_edits():
    """Sphere-add r=64 (centre seed 1), then sphere-remove r=48 (centre seed 2), in [0,1024)^3."""
    c1 = np.random.default_rng(1).uniform(0, 1024, 3).astype(np.float32)
    c2 = np.random.default_rng(2).uniform(0, 1024, 3).astype(np.float32)
    return [(SPHERE_ADD, 2, tuple(float(x) for x in c1), 64.0),
            (SPHERE_REMOVE, 0, tuple(float(x) for x in c2), 48.0)]


def surface_edits(heights, n, seed, levels, rmin=8.0, rmax=512.0, kinds=(0.4, 0.4, 0.2), near=1024.0):
    """Edit stream near the terrain surface: sphere-add / sphere-remove / box-fill
    with the given probabilities, radius log-uniform in [rmin, rmax], centre
    (x, h(x,z) + U[-near, near], z)."""
    rng = np.random.default_rng(seed)
    side = 1 << levels
    out = []
    for _ in range(n):
        k = rng.choice(3, p=kinds)
        kind = (SPHERE_ADD, SPHERE_REMOVE, BOX_ADD)[k]
        r = float(np.exp(rng.uniform(np.log(rmin), np.log(rmax))))
        x, z = rng.integers(0, side, 2)
        y = float(heights(int(x), int(z))) + float(rng.uniform(-near, near))
        y = min(max(y, 0.0), side - 1.0)
        mat = 0 if kind == SPHERE_REMOVE else int(rng.integers(2, 15))
        out.append((kind, mat, (float(x), y, float(z)), r))
    return out


def config1_camera(levels=10):
    side = float(1 << levels)
    return dict(position=(-0.1 * side, 0.85 * side, -0.1 * side), target=(0.5 * side, 0.35 * side, 0.5 * side),
                fov=60.0, bmin=(0.0, 0.0, 0.0), bmax=(side, side, side))


def voxel_edit_count(kind, center, radius, levels):
    """Voxels whose unit box satisfies the tool's overlap predicate, by per-row
    interval counting (the unit of "voxel edits/s").  Spheres use the tool's
    centre + 0.5 and radius >= 0.1; boxes use [c - r, c + r) (r = 0 -> one voxel)."""
    side = 1 << levels
    if kind in (SPHERE_ADD, SPHERE_REMOVE, PAINT):
        c = np.asarray(center, np.float64) + 0.5
        R = max(float(radius), 0.1)
        ys = np.arange(max(0, int(np.floor(c[1] - R)) - 1), min(side, int(np.ceil(c[1] + R)) + 1))
        zs = np.arange(max(0, int(np.floor(c[2] - R)) - 1), min(side, int(np.ceil(c[2] + R)) + 1))
        if len(ys) == 0 or len(zs) == 0:
            return 0

        def gap(cc, v):
            return np.where(cc < v, v - cc, np.where(cc > v + 1, cc - (v + 1), 0.0))

        gy = gap(c[1], ys.astype(np.float64))[:, None]
        gz = gap(c[2], zs.astype(np.float64))[None, :]
        T = R * R - gy * gy - gz * gz
        s = np.sqrt(np.maximum(T, 0.0))
        # x overlaps iff the gap from c.x to [x, x+1] is < s: x in (c.x - s - 1, c.x + s).
        lo = np.maximum(np.floor(c[0] - s - 1.0) + 1.0, 0.0)
        hi = np.minimum(np.ceil(c[0] + s) - 1.0, side - 1.0)
        n = np.where(T > 0, np.maximum(hi - lo + 1.0, 0.0), 0.0)
        return int(n.sum())
    r = float(radius)
    total = 1
    for a in range(3):
        bmin = center[a] - r
        bmax = center[a] + (1.0 if r == 0 else r)
        lo = max(int(np.floor(bmin - 1.0)) + 1, 0)
        hi = min(int(np.ceil(bmax)) - 1, side - 1)
        total *= max(hi - lo + 1, 0)
    return total


def config3_edits(height_at, levels, n=256, seed=6):
    """Edit stream of the batched-editing workload: 40% sphere-add, 40%
    sphere-remove, 20% axis-aligned box fill; radius log-uniform in [8, 512];
    centre (x, h(x,z) + U[-1024, 1024], z), x, z ~ U[0, 2^D)."""
    return surface_edits(height_at, n, seed, levels, rmin=8.0, rmax=512.0, kinds=(0.4, 0.4, 0.2), near=1024.0)


def config4_requests(height_at, levels, rate, duration_s, seed=7):
    """Serving workload: Poisson arrivals at `rate` per second for
    `duration_s`; the radius / centre mix of the batched-editing workload
    (log-uniform r in [8, 512], centre within +-1024 of the surface) over
    sphere-add 35%, sphere-remove 35%, box fill 15% and a sphere paint 15%
    that sets a 4-bit material id.  Returns (edits, arrival seconds)."""
    rng = np.random.default_rng(seed)
    side = 1 << levels
    arrivals = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration_s:
            break
        arrivals.append(t)
    edits = []
    for _ in arrivals:
        k = int(rng.choice(4, p=(0.35, 0.35, 0.15, 0.15)))
        kind = (SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, PAINT)[k]
        r = float(np.exp(rng.uniform(np.log(8.0), np.log(512.0))))
        x, z = rng.integers(0, side, 2)
        y = float(height_at(int(x), int(z))) + float(rng.uniform(-1024.0, 1024.0))
        y = min(max(y, 0.0), side - 1.0)
        mat = 0 if kind == SPHERE_REMOVE else int(rng.integers(1, 16))
        edits.append((kind, mat, (float(x), y, float(z)), r))
    return edits, np.asarray(arrivals, np.float64)


def camera_path(levels, n, seed=8):
    """Seeded fly-over: n cameras on a circle above the terrain looking at
    points near the centre, as (position, target, fov, bmin, bmax)."""
    rng = np.random.default_rng(seed)
    side = float(1 << levels)
    out = []
    phase = float(rng.uniform(0, 2 * np.pi))
    for i in range(n):
        a = phase + 2 * np.pi * i / max(n, 1)
        pos = (0.5 * side + 0.6 * side * np.cos(a), 0.9 * side, 0.5 * side + 0.6 * side * np.sin(a))
        tgt = (0.5 * side + float(rng.uniform(-0.05, 0.05)) * side, 0.45 * side,
               0.5 * side + float(rng.uniform(-0.05, 0.05)) * side)
        out.append((pos, tgt, 60.0, (0.0, 0.0, 0.0), (side, side, side)))
    return out


def band_materials(x, y, z, material, band_log2, band_seed, material_bits):
    """Material of solid voxels under svdag_build_terrain_banded's height-band
    rule, vectorised: (material + (y >> b) + (mix(seed ^ (x>>b)*0x9E3779B1 ^
    (z>>b)*0x85EBCA77) & 3)) & (2^M - 1), mix(h) = t ^ (t >> 12) with
    t = (h ^ h >> 15) * 0x2C1B3C6D, all in u32 arithmetic."""
    x, y, z = (np.asarray(v, np.uint64) for v in (x, y, z))
    b = np.uint64(band_log2)
    m32 = np.uint64(0xFFFFFFFF)
    h = (np.uint64(band_seed) ^ (((x >> b) * np.uint64(0x9E3779B1)) & m32)
         ^ (((z >> b) * np.uint64(0x85EBCA77)) & m32))
    h = ((h ^ (h >> np.uint64(15))) * np.uint64(0x2C1B3C6D)) & m32
    h ^= h >> np.uint64(12)
    mask = np.uint64((1 << material_bits) - 1)
    return ((np.uint64(material) + (y >> b) + (h & np.uint64(3))) & mask).astype(np.uint32)


def copy_box_voxels(center, radius, n, rng, levels):
    """n seeded destination voxels of a copy whose region is
    [center - radius, center + radius): voxels whose unit box overlaps it."""
    lo = [max(int(np.floor(c - radius - 1.0)) + 1, 0) for c in center]
    hi = [min(int(np.ceil(c + radius)) - 1, (1 << levels) - 1) for c in center]
    return np.stack([rng.integers(lo[a], hi[a] + 1, n) for a in range(3)], axis=1).astype(np.int64)


def copy_source_voxels(v, center, radius, src, dst, rotation):
    """Source voxel each destination voxel v of a copy/paste reads, as the
    COPY tool defines it: c = dst + (center - src), pivot p = floor(c + 0.5),
    q = p + trunc(src - dst), source = max(q + R^k(v - p), 0) with
    R(x, y, z) = (-z - 1, y, x) (k quarter turns about +Y), in float32 for
    the centre arithmetic as the tool computes it."""
    f = np.float32
    c = np.array([f(dst[a]) + (f(center[a]) - f(src[a])) for a in range(3)], np.float32)
    p = np.floor(c + f(0.5)).astype(np.int64)
    q = p + np.trunc(np.array([f(src[a]) - f(dst[a]) for a in range(3)], np.float32)).astype(np.int64)
    d = np.asarray(v, np.int64) - p
    for _ in range(int(rotation) & 3):
        d = np.stack([-d[:, 2] - 1, d[:, 1], d[:, 0]], axis=1)
    return np.maximum(q + d, 0)
