"""ctypes loader of liboracle.so -- test infrastructure only (tests/, smoke(),
bench.py's cpu_baseline leg).  Never imported by the product."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.ora_table_create.restype = ctypes.c_void_p
        L.ora_table_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.ora_table_destroy.argtypes = [ctypes.c_void_p]
        for f in ("ora_table_add", "ora_table_decode"):
            getattr(L, f).argtypes = [ctypes.c_void_p, u32p, ctypes.c_uint32, u32p]
        L.ora_table_find.argtypes = [ctypes.c_void_p, u32p, ctypes.c_uint32, u32p, ctypes.c_int]
        L.ora_xor_hash.restype = ctypes.c_uint32
        L.ora_xor_hash.argtypes = [u32p, ctypes.c_uint32]
        L.ora_accel_hash.restype = ctypes.c_uint32
        L.ora_accel_hash.argtypes = [u32p, ctypes.c_uint32]
        L.ora_table_overflow_slabs.restype = ctypes.c_uint32
        L.ora_table_overflow_slabs.argtypes = [ctypes.c_void_p]
        L.ora_table_raw.argtypes = [ctypes.c_void_p, u32p, u32p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(u32p)


class OracleTable:
    """The reference CompactAccelerationHashTable (host path), restated."""

    def __init__(self, item_words, num_buckets):
        self.s = item_words
        self.h = lib().ora_table_create(item_words, num_buckets)

    def add(self, items):
        items = np.ascontiguousarray(items, dtype=np.uint32).reshape(-1, self.s)
        out = np.empty(len(items), np.uint32)
        lib().ora_table_add(self.h, _p(items), len(items), _p(out))
        return out

    def find(self, items, threads=1):
        items = np.ascontiguousarray(items, dtype=np.uint32).reshape(-1, self.s)
        out = np.empty(len(items), np.uint32)
        lib().ora_table_find(self.h, _p(items), len(items), _p(out), threads)
        return out

    def decode(self, ptrs):
        ptrs = np.ascontiguousarray(ptrs, dtype=np.uint32)
        out = np.empty((len(ptrs), self.s), np.uint32)
        lib().ora_table_decode(self.h, _p(ptrs), len(ptrs), _p(out))
        return out

    def overflow_slabs(self):
        return lib().ora_table_overflow_slabs(self.h)

    def raw(self, num_buckets):
        """(head slabs, overflow slabs) as flat u32 arrays."""
        sw = 10 + 32 * self.s
        heads = np.empty(num_buckets * sw, np.uint32)
        pool = np.empty(self.overflow_slabs() * sw, np.uint32)
        lib().ora_table_raw(self.h, _p(heads), _p(pool))
        return heads, pool

    def __del__(self):
        try:
            lib().ora_table_destroy(self.h)
        except Exception:
            pass


def xor_hash(item):
    a = np.ascontiguousarray(item, dtype=np.uint32)
    return lib().ora_xor_hash(_p(a), len(a))


def accel_hash(item):
    a = np.ascontiguousarray(item, dtype=np.uint32)
    return lib().ora_accel_hash(_p(a), len(a))


def _dag_sigs(L):
    vp = ctypes.c_void_p
    f32p = ctypes.POINTER(ctypes.c_float)
    f64p = ctypes.POINTER(ctypes.c_double)
    szp = ctypes.POINTER(ctypes.c_size_t)
    L.ora_dag_create.restype = vp
    L.ora_dag_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.ora_dag_destroy.argtypes = [vp]
    L.ora_dag_root.restype = ctypes.c_uint32
    L.ora_dag_root.argtypes = [vp]
    L.ora_dag_set_root.argtypes = [vp, ctypes.c_uint32]
    L.ora_dag_item_count.restype = ctypes.c_uint64
    L.ora_dag_item_count.argtypes = [vp]
    L.ora_terrain_heights.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]
    L.ora_dag_build_terrain.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.ora_dag_edit.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, f32p, ctypes.c_float, f32p, f32p, ctypes.c_uint32,
                               u32p]
    L.ora_dag_query.argtypes = [vp, u32p, ctypes.c_uint32, u32p]
    L.ora_dag_canonical.restype = ctypes.c_uint64
    L.ora_dag_canonical.argtypes = [vp, u32p, szp, u32p]
    L.ora_canonical_from_records.restype = ctypes.c_uint64
    L.ora_canonical_from_records.argtypes = [u32p, ctypes.c_size_t, u32p, szp, u32p]
    L.ora_dag_live_items.restype = ctypes.c_uint64
    L.ora_dag_live_items.argtypes = [vp]
    L.ora_dag_trace.argtypes = [vp, f64p, f64p, ctypes.c_double, f64p, f64p, ctypes.c_uint32, ctypes.c_uint32, u32p]
    L.ora_dag_shade.argtypes = [vp, f64p, f64p, ctypes.c_double, f64p, f64p, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_float, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, f32p, f32p]
    L.ora_dag_occluded.restype = ctypes.c_int
    L.ora_dag_occluded.argtypes = [vp, f32p, f32p, ctypes.c_float]
    L.ora_dag_ao_at.restype = ctypes.c_float
    L.ora_dag_ao_at.argtypes = [vp, f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                ctypes.c_uint64]
    L.ora_pcg32_raw.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, u32p]
    L.ora_pcg32_seeded.argtypes = [ctypes.c_uint64, ctypes.c_uint32, u32p]
    L.ora_sincos_turns.argtypes = [ctypes.c_float, f32p, f32p]
    L.ora_sun_dir.argtypes = [f32p]
    L.ora_leaf_build.restype = ctypes.c_uint32
    L.ora_leaf_build.argtypes = [ctypes.c_uint32, u32p, u32p, u32p]
    L.ora_leaf_material.argtypes = [ctypes.c_uint32, u32p, ctypes.c_uint32, u32p]
    L.ora_leaf_size.restype = ctypes.c_uint32
    L.ora_leaf_size.argtypes = [ctypes.c_uint32, u32p]
    L.ora_dag_build_dense.restype = ctypes.c_uint32
    L.ora_dag_build_dense.argtypes = [vp, ctypes.c_void_p]
    L.ora_header_make.restype = ctypes.c_uint32
    L.ora_header_make.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.ora_header_decode.argtypes = [ctypes.c_uint32, ctypes.c_uint32] + [u32p] * 4
    L.ora_sphere_overlap.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_float,
                                     ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.ora_dag_path_trace.argtypes = [vp, f64p, f64p, ctypes.c_double, f64p, f64p, ctypes.c_uint32, ctypes.c_uint32,
                                     u32p, f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_uint64, f32p, ctypes.c_uint32, u32p]
    L.ora_srgb_to_linear.restype = ctypes.c_float
    L.ora_srgb_to_linear.argtypes = [ctypes.c_float]
    L.ora_linear_to_srgb.restype = ctypes.c_float
    L.ora_linear_to_srgb.argtypes = [ctypes.c_float]
    L.ora_rgb888.restype = ctypes.c_uint32
    L.ora_rgb888.argtypes = [f32p]
    L.ora_blur_weights.argtypes = [f32p]
    L.ora_colors.argtypes = [u32p, ctypes.c_uint32, f32p, ctypes.c_uint32, u32p]
    L.ora_blur.argtypes = [f32p, ctypes.c_uint32, ctypes.c_uint32, f32p]
    L.ora_lighting.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, f32p, f32p, f64p, f64p, ctypes.c_double,
                               f64p, f64p, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, u32p]


_dag_ready = False


def dlib():
    global _dag_ready
    L = lib()
    if not _dag_ready:
        _dag_sigs(L)
        _dag_ready = True
    return L


SENT = 0xFFFFFFFF
HIT_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("material", "<u4"),
                      ("t", "<f4"), ("px", "<f4"), ("py", "<f4"), ("pz", "<f4")])


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def _d(v):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


def terrain_heights(seed, base, amp, freq, octaves, x0, z0, w, d):
    out = np.empty((d, w), np.int32)
    dlib().ora_terrain_heights(seed, base, amp, freq, octaves, x0, z0, w, d,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out


class OracleDag:
    """The reference's GPU editing path, restated on the CPU."""

    def __init__(self, levels, material_bits=4):
        self.levels = levels
        self.h = dlib().ora_dag_create(levels, material_bits)
        if not self.h:
            raise ValueError("bad oracle DAG parameters")

    def build_terrain(self, seed, base, amp, freq, octaves, material, band_log2=0, band_seed=0):
        dlib().ora_dag_build_terrain(self.h, seed, base, amp, freq, octaves, material, band_log2, band_seed)

    @property
    def root(self):
        return dlib().ora_dag_root(self.h)

    @root.setter
    def root(self, r):
        dlib().ora_dag_set_root(self.h, r)

    def edit(self, kind, material, center, radius, src=(0, 0, 0), dst=(0, 0, 0), rotation=0):
        counts = (ctypes.c_uint32 * 32)()
        applied = dlib().ora_dag_edit(self.h, kind, material, _f3(center), float(radius), _f3(src), _f3(dst), rotation,
                                      counts)
        return applied, list(counts)

    def query(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32).reshape(-1, 3)
        out = np.empty(len(xyz), np.uint32)
        dlib().ora_dag_query(self.h, _p(xyz), len(xyz), _p(out))
        return out

    def canonical(self, dump=False):
        n = ctypes.c_size_t(0)
        counts = (ctypes.c_uint32 * 32)()
        h = dlib().ora_dag_canonical(self.h, None, ctypes.byref(n), counts)
        if not dump:
            return h, list(counts)
        buf = np.empty(n.value, np.uint32)
        dlib().ora_dag_canonical(self.h, _p(buf), ctypes.byref(n), counts)
        return h, list(counts), buf

    def live_items(self):
        return dlib().ora_dag_live_items(self.h)

    def trace(self, position, rotation, fov, bmin, bmax, w, h):
        out = np.empty(w * h, HIT_DTYPE)
        dlib().ora_dag_trace(self.h, _d(position), _d(rotation), float(fov), _d(bmin), _d(bmax), w, h,
                             out.ctypes.data_as(u32p))
        return out

    def shade(self, position, rotation, fov, bmin, bmax, w, h, bias=1.0, ao_samples=3, ao_length=32.0, seed=1):
        """(sun visibility * max(N.L, 0), ambient occlusion) per pixel, stored layout."""
        sun = np.empty(w * h, np.float32)
        ao = np.empty(w * h, np.float32)
        f32p = ctypes.POINTER(ctypes.c_float)
        dlib().ora_dag_shade(self.h, _d(position), _d(rotation), float(fov), _d(bmin), _d(bmax), w, h, float(bias),
                             ao_samples, float(ao_length), seed, sun.ctypes.data_as(f32p), ao.ctypes.data_as(f32p))
        return sun, ao

    def path_trace(self, position, rotation, fov, bmin, bmax, w, h, colors_rgba8, palette, env=1.0, samples=1,
                   depth=2, integrate=False, seed=1, accum=None, accumulated=0):
        """trace_path_tracing per stored pixel; returns (RGBA8 image, accumulation sums)."""
        f32p = ctypes.POINTER(ctypes.c_float)
        col = np.ascontiguousarray(colors_rgba8 if colors_rgba8 is not None else np.zeros(w * h), np.uint32)
        pal = np.ascontiguousarray(palette, np.float32).reshape(-1, 3)
        acc = np.zeros(3 * w * h, np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        out = np.empty(w * h, np.uint32)
        dlib().ora_dag_path_trace(self.h, _d(position), _d(rotation), float(fov), _d(bmin), _d(bmax), w, h, _p(col),
                                  pal.ctypes.data_as(f32p), len(pal), float(env), samples, depth, int(bool(integrate)),
                                  seed, acc.ctypes.data_as(f32p), accumulated, _p(out))
        return out, acc

    def occluded(self, origin, direction, tmax=3.4e38):
        return bool(dlib().ora_dag_occluded(self.h, _f3(origin), _f3(direction), float(tmax)))

    def ao_at(self, point, code, samples, bias, length, stream):
        return float(dlib().ora_dag_ao_at(self.h, _f3(point), code, samples, float(bias), float(length), stream))

    def __del__(self):
        try:
            dlib().ora_dag_destroy(self.h)
        except Exception:
            pass


def canonical_from_records(records):
    records = np.ascontiguousarray(records, dtype=np.uint32)
    n = ctypes.c_size_t(0)
    counts = (ctypes.c_uint32 * 32)()
    h = dlib().ora_canonical_from_records(_p(records), len(records), None, ctypes.byref(n), counts)
    return h, list(counts)


def leaf_build(material_bits, present, materials):
    present = np.ascontiguousarray(present, dtype=np.uint32)
    materials = np.ascontiguousarray(materials, dtype=np.uint32)
    out = np.zeros(20, np.uint32)
    n = dlib().ora_leaf_build(material_bits, _p(present), _p(materials), _p(out))
    return out[:n]


def leaf_material(material_bits, leaf, voxel):
    leaf = np.ascontiguousarray(leaf, dtype=np.uint32)
    m = ctypes.c_uint32(0)
    ok = dlib().ora_leaf_material(material_bits, _p(leaf), voxel, ctypes.byref(m))
    return m.value if ok else None


def leaf_size(material_bits, leaf):
    leaf = np.ascontiguousarray(leaf, dtype=np.uint32)
    return dlib().ora_leaf_size(material_bits, _p(leaf))


def header_make(material_bits, mask, payload):
    return dlib().ora_header_make(material_bits, mask, payload)


def header_decode(material_bits, header):
    vals = [ctypes.c_uint32() for _ in range(4)]
    dlib().ora_header_decode(material_bits, header, *[ctypes.cast(ctypes.byref(v), u32p) for v in vals])
    return tuple(v.value for v in vals)


def sphere_overlap(center, radius, mn, mx):
    L = dlib()
    f3 = ctypes.c_float * 3
    return bool(L.ora_sphere_overlap(f3(*center), ctypes.c_float(radius), f3(*mn), f3(*mx)))


def build_dense(dag, vox):
    """Factory-rule DAG over a dense voxel grid vox[x, y, z] (uint8, 0xFF = empty)."""
    vox = np.ascontiguousarray(vox, dtype=np.uint8)
    dlib().ora_dag_build_dense(dag.h, vox.ctypes.data_as(ctypes.c_void_p))


def pcg32_raw(state, inc, n):
    out = np.empty(n, np.uint32)
    dlib().ora_pcg32_raw(state, inc, n, out.ctypes.data_as(u32p))
    return out


def pcg32_seeded(stream, n):
    out = np.empty(n, np.uint32)
    dlib().ora_pcg32_seeded(stream, n, out.ctypes.data_as(u32p))
    return out


def sincos_turns(u):
    sn, cs = ctypes.c_float(0), ctypes.c_float(0)
    dlib().ora_sincos_turns(ctypes.c_float(u), ctypes.byref(sn), ctypes.byref(cs))
    return sn.value, cs.value


def sun_dir():
    L = (ctypes.c_float * 3)()
    dlib().ora_sun_dir(L)
    return np.array(list(L), np.float32)


_hd_ready = False


def hdlib():
    """liboracle.so with the CPU HashDAG entry points typed."""
    global _hd_ready
    L = lib()
    if not _hd_ready:
        L.hd_create.restype = ctypes.c_void_p
        L.hd_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.hd_destroy.argtypes = [ctypes.c_void_p]
        L.hd_import_canonical.restype = ctypes.c_int
        L.hd_import_canonical.argtypes = [ctypes.c_void_p, u32p, ctypes.c_uint64, ctypes.c_uint32]
        L.hd_edit.restype = ctypes.c_int
        L.hd_edit.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_uint64)]
        L.hd_query.argtypes = [ctypes.c_void_p, u32p, ctypes.c_uint64, ctypes.c_void_p]
        L.hd_usage.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        _hd_ready = True
    return L


class CpuHashDag:
    """The reference's CPU HashDAG editing path (geometry only), restated:
    the CPU timing baseline.  kinds: 0 sphere add, 1 sphere remove, 2 box
    add, 3 box remove."""

    SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, BOX_REMOVE = range(4)

    def __init__(self, levels, color_levels=0, top_span=0, low_span=0, pool_bytes=64 << 30):
        self.levels = levels
        self.h = hdlib().hd_create(levels, color_levels, top_span, low_span, pool_bytes)
        if not self.h:
            raise RuntimeError("hd_create failed (bad levels / bucket spans, or no address space for the pool)")

    def import_canonical(self, dump, threads=1):
        d = np.ascontiguousarray(dump, np.uint32)
        rc = hdlib().hd_import_canonical(self.h, _p(d), d.size, threads)
        if rc != 0:
            raise RuntimeError("hd_import_canonical: " + ("malformed dump" if rc == -1 else "a bucket overflowed"))

    def edit(self, kind, center, radius, threads=32):
        """Returns (wall ms of the edit, leaf tasks)."""
        ms, tasks = ctypes.c_double(), ctypes.c_uint64()
        rc = hdlib().hd_edit(self.h, kind, float(center[0]), float(center[1]), float(center[2]), float(radius),
                             threads, ctypes.byref(ms), ctypes.byref(tasks))
        if rc != 0:
            raise RuntimeError("hd_edit: a bucket or the pool overflowed")
        return ms.value, tasks.value

    def query(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
        out = np.empty(len(xyz), np.uint8)
        hdlib().hd_query(self.h, _p(xyz), len(xyz), out.ctypes.data_as(ctypes.c_void_p))
        return out

    def usage(self):
        out = (ctypes.c_uint64 * 2)()
        hdlib().hd_usage(self.h, out)
        return int(out[0]), int(out[1])

    def __del__(self):
        if getattr(self, "h", None):
            hdlib().hd_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- colour, blur, lighting (oracle_shading.cpp)
_f32p = ctypes.POINTER(ctypes.c_float)


def srgb_to_linear(c):
    return dlib().ora_srgb_to_linear(float(c))


def linear_to_srgb(c):
    return dlib().ora_linear_to_srgb(float(c))


def rgb888(c):
    a = np.ascontiguousarray(c, np.float32)
    return dlib().ora_rgb888(a.ctypes.data_as(_f32p))


def blur_weights():
    w = np.empty(9, np.float32)
    dlib().ora_blur_weights(w.ctypes.data_as(_f32p))
    return w


def _hits_u32(hits):
    return np.ascontiguousarray(hits).view(np.uint32).reshape(-1)


def colors(hits, palette):
    """RGBA8 diffuse colour per stored pixel (trace_colors with one-colour materials)."""
    h = _hits_u32(hits)
    pal = np.ascontiguousarray(palette, np.float32).reshape(-1, 3)
    out = np.empty(len(h) // 8, np.uint32)
    dlib().ora_colors(_p(h), len(out), pal.ctypes.data_as(_f32p), len(pal), _p(out))
    return out


def blur(values, w, h):
    v = np.ascontiguousarray(values, np.float32)
    out = np.empty(w * h, np.float32)
    dlib().ora_blur(v.ctypes.data_as(_f32p), w, h, out.ctypes.data_as(_f32p))
    return out


def lighting(hits, w, h, colors_rgba8, ao, sun, position, rotation, fov, bmin, bmax, levels, sun_brightness=0.0,
             fog_density=0.0):
    """Final RGBA8 per stored pixel (trace_lighting); ao / sun None = not applied."""
    hh = _hits_u32(hits)
    col = np.ascontiguousarray(colors_rgba8, np.uint32)
    a = None if ao is None else np.ascontiguousarray(ao, np.float32)
    s = None if sun is None else np.ascontiguousarray(sun, np.float32)
    out = np.empty(w * h, np.uint32)
    dlib().ora_lighting(_p(hh), w, h, _p(col), None if a is None else a.ctypes.data_as(_f32p),
                        None if s is None else s.ctypes.data_as(_f32p), _d(position), _d(rotation), float(fov),
                        _d(bmin), _d(bmax), levels, float(sun_brightness), float(fog_density), _p(out))
    return out
