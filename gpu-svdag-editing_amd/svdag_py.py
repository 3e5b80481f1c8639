"""ctypes binding of libsvdag (include/svdag.h) used by the tests and bench.py.

Loads gpu-svdag-editing_amd/lib/libsvdag.so and fails loudly if it is missing:
there is no fallback path.  The classes mirror the reference's interface
names (MyGPUHashDAG, MyGPUHashDAGUndoRedo, DAGTracer) over the C ABI.
"""
import ctypes
import struct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsvdag.so")

SVDAG_OK = 0
ERRORS = {-1: "INVALID", -2: "OOM", -3: "DEVICE", -4: "CAPACITY", -5: "EMPTY", -6: "NO_DEVICE"}

SPHERE_ADD, SPHERE_REMOVE, BOX_ADD, BOX_REMOVE, PAINT, COPY = range(6)
NOT_FOUND = 0xFFFFFFFF
QUERY_NOALLOC_MAX = 1 << 20  # points per svdag_query_voxels_noalloc call

u32p = ctypes.POINTER(ctypes.c_uint32)


class Lighting(ctypes.Structure):
    _fields_ = [("apply_ambient_occlusion", ctypes.c_uint32), ("apply_shadows", ctypes.c_uint32),
                ("fog_density", ctypes.c_float), ("sun_brightness", ctypes.c_float)]


class PathTracing(ctypes.Structure):
    _fields_ = [("environment_brightness", ctypes.c_float), ("num_samples", ctypes.c_uint32),
                ("max_path_depth", ctypes.c_uint32), ("integrate_pixel", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("accumulate", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class SvdagError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"svdag error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Segment(ctypes.Structure):
    _fields_ = [("item_words", ctypes.c_uint32), ("count", ctypes.c_uint32),
                ("word_offset", ctypes.c_uint32), ("out_offset", ctypes.c_uint32)]


class Config(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("material_bits", ctypes.c_uint32),
                ("load_factor", ctypes.c_uint32), ("group_size", ctypes.c_uint32),
                ("expected_items", ctypes.c_uint64), ("arena_slabs", ctypes.c_uint64),
                ("workspace_items", ctypes.c_uint64)]


class Terrain(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("base", ctypes.c_double), ("amplitude", ctypes.c_double),
                ("frequency", ctypes.c_double), ("octaves", ctypes.c_uint32), ("material", ctypes.c_uint32)]


class EditDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("material", ctypes.c_uint32), ("center", ctypes.c_float * 3),
                ("radius", ctypes.c_float), ("src", ctypes.c_float * 3), ("dst", ctypes.c_float * 3),
                ("rotation", ctypes.c_uint32)]


_EDIT_PACK = struct.Struct("<II3ff3f3fI")
assert _EDIT_PACK.size == ctypes.sizeof(EditDesc)


class EditStats(ctypes.Structure):
    _fields_ = [("svo_items", ctypes.c_uint32 * 32), ("applied", ctypes.c_uint32),
                ("root_before", ctypes.c_uint32), ("root_after", ctypes.c_uint32), ("gpu_ms", ctypes.c_float)]


class GcStats(ctypes.Structure):
    _fields_ = [("items_freed", ctypes.c_uint64), ("slabs_freed", ctypes.c_uint64),
                ("items_live", ctypes.c_uint64), ("gpu_ms", ctypes.c_float), ("marked", ctypes.c_uint32)]


class Camera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 3), ("rotation", ctypes.c_double * 9),
                ("fov_degrees", ctypes.c_double), ("bounds_min", ctypes.c_double * 3),
                ("bounds_max", ctypes.c_double * 3)]


class TuneResult(ctypes.Structure):
    _fields_ = [("group_size", ctypes.c_uint32), ("ms", ctypes.c_float * 4)]


class TuneReport(ctypes.Structure):
    _fields_ = [("group_size", ctypes.c_uint32), ("max_blocks", ctypes.c_uint32), ("ms_group", ctypes.c_float * 4),
                ("ms_blocks", ctypes.c_float * 5)]


class Memory(ctypes.Structure):
    _fields_ = [("slab_capacity", ctypes.c_uint64), ("slabs_used", ctypes.c_uint64), ("slabs_free", ctypes.c_uint64),
                ("bytes_reserved", ctypes.c_uint64), ("min_free_fraction", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("frame_width", ctypes.c_uint32), ("frame_height", ctypes.c_uint32),
                ("frame_period_us", ctypes.c_uint32), ("cameras", ctypes.POINTER(Camera)),
                ("num_cameras", ctypes.c_uint32), ("allow_gc", ctypes.c_uint32), ("reserve_fraction", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("frames", ctypes.c_uint64),
                ("frames_skipped", ctypes.c_uint64), ("gcs", ctypes.c_uint64), ("grows", ctypes.c_uint64),
                ("deferred", ctypes.c_uint64), ("wall_s", ctypes.c_double), ("edit_busy_s", ctypes.c_double),
                ("trace_ms_avg", ctypes.c_double), ("last_frame_root", ctypes.c_uint32),
                ("max_batch_seen", ctypes.c_uint32), ("sequential_spine_edits", ctypes.c_uint64),
                ("handed_back_edits", ctypes.c_uint64)]


REQUEST_DTYPE = np.dtype([("status", "<i4"), ("applied", "<u4"), ("root_after", "<u4"), ("batch", "<u4"),
                          ("submit_s", "<f8"), ("admit_s", "<f8"), ("done_s", "<f8")])

HIT_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("material", "<u4"),
                      ("t", "<f4"), ("px", "<f4"), ("py", "<f4"), ("pz", "<f4")])

# Every symbol include/svdag.h declares (checked by the CPU test suite).
EXPORTED = [
    "svdag_last_error", "svdag_abi_version", "svdag_device_count", "svdag_dev_alloc", "svdag_dev_free",
    "svdag_memcpy_h2d", "svdag_memcpy_d2h", "svdag_memcpy_d2d", "svdag_device_synchronize",
    "svdag_hashset_create", "svdag_hashset_destroy", "svdag_hashset_add_dev", "svdag_hashset_find_dev",
    "svdag_hashset_read", "svdag_hashset_usage",
    "svdag_create", "svdag_destroy", "svdag_build_terrain", "svdag_build_terrain_banded", "svdag_terrain_heights", "svdag_root",
    "svdag_edit", "svdag_edit_batch", "svdag_edit_stream", "svdag_hashset_timer", "svdag_autotune", "svdag_set_batch_mode", "svdag_undo", "svdag_redo", "svdag_add_frame", "svdag_history", "svdag_gc", "svdag_gc_mark", "svdag_query_voxels",
    "svdag_query_voxels_noalloc", "svdag_edit_on_stream", "svdag_autotune_all",
    "svdag_trace_primary", "svdag_memory_usage", "svdag_server_start", "svdag_server_submit", "svdag_server_wait",
    "svdag_server_stop", "svdag_server_last_frame", "svdag_serve_replay", "svdag_export_canonical",
    "svdag_export_tables", "svdag_import_canonical", "svdag_save", "svdag_load", "svdag_hbm_probe",
    "svdag_trace_shadows", "svdag_trace_ambient_occlusion", "svdag_build_voxels",
    "svdag_read_reference_dag", "svdag_write_reference_dag", "svdag_read_reference_scene",
    "svdag_write_reference_scene", "svdag_set_load_factor", "svdag_autotune_load_factor",
    "svdag_blur_ambient_occlusion", "svdag_trace_colors", "svdag_trace_lighting", "svdag_trace_path_tracing",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with gpu-svdag-editing_amd/build_lib.py")
        L = ctypes.CDLL(LIB_PATH)
        L.svdag_last_error.restype = ctypes.c_char_p
        L.svdag_abi_version.restype = ctypes.c_uint32
        _lib = L
    return _lib


def check(code):
    if code != SVDAG_OK:
        raise SvdagError(code, lib().svdag_last_error().decode(errors="replace"))


def device_count():
    n = ctypes.c_int(0)
    check(lib().svdag_device_count(ctypes.byref(n)))
    return n.value


PROBE_STREAM, PROBE_RANDOM_LINES = 0, 1


def hbm_probe(kind, nbytes, reads=0):
    """Best kernel ms of a stream read of nbytes (PROBE_STREAM) or of `reads`
    random line-head reads (a 128-byte line's first 32 bytes) in an nbytes
    buffer (PROBE_RANDOM_LINES)."""
    ms = ctypes.c_float(0.0)
    check(lib().svdag_hbm_probe(ctypes.c_uint32(kind), ctypes.c_uint64(nbytes), ctypes.c_uint64(reads),
                                ctypes.byref(ms)))
    return ms.value


class DeviceBuffer:
    """Device allocation owned by the caller (plumbing only)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        check(lib().svdag_dev_alloc(ctypes.c_size_t(self.nbytes), ctypes.byref(self.ptr)))

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().svdag_memcpy_h2d(b.ptr, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes)))
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().svdag_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes)))
        return out

    def free(self):
        if self.ptr:
            check(lib().svdag_dev_free(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_segments(groups):
    """groups: list of (item_words, uint32 array [count, item_words]) -> (flat words, Segment array)."""
    segs = (Segment * len(groups))()
    flat = []
    woff = 0
    ooff = 0
    for i, (s, arr) in enumerate(groups):
        arr = np.ascontiguousarray(arr, dtype=np.uint32).reshape(-1, s)
        segs[i].item_words = s
        segs[i].count = arr.shape[0]
        segs[i].word_offset = woff
        segs[i].out_offset = ooff
        flat.append(arr.reshape(-1))
        woff += arr.size
        ooff += arr.shape[0]
    words = np.concatenate(flat) if flat else np.zeros(0, np.uint32)
    return words, segs, ooff


class HashSet:
    """Slab hash tables, one per item size (CompactAccelerationHashTable per size)."""

    def __init__(self, min_words, max_words, buckets, reserve=None):
        n = max_words - min_words + 1
        b = (ctypes.c_uint32 * n)(*[int(x) for x in buckets])
        r = (ctypes.c_uint32 * n)(*[int(x) for x in (reserve if reserve is not None else [0] * n)])
        self.h = ctypes.c_void_p()
        check(lib().svdag_hashset_create(min_words, max_words, b, r, ctypes.byref(self.h)))
        self.min_words, self.max_words = min_words, max_words

    def _bulk(self, fn, words_dev, segs, out_dev, variant):
        ms = ctypes.c_float(0)
        check(fn(self.h, ctypes.cast(words_dev.ptr, u32p), segs, len(segs), ctypes.cast(out_dev.ptr, u32p),
                 int(variant), ctypes.byref(ms)))
        return ms.value

    def add_dev(self, words_dev, segs, out_dev, variant=16):
        return self._bulk(lib().svdag_hashset_add_dev, words_dev, segs, out_dev, variant)

    def find_dev(self, words_dev, segs, out_dev, variant=16, timed=True):
        """timed=False (variants below 1000): only queue the find; returns None."""
        if not timed:
            check(lib().svdag_hashset_find_dev(self.h, ctypes.cast(words_dev.ptr, u32p), segs, len(segs),
                                               ctypes.cast(out_dev.ptr, u32p), int(variant), None))
            return None
        return self._bulk(lib().svdag_hashset_find_dev, words_dev, segs, out_dev, variant)

    def timer(self, stop):
        """svdag_hashset_timer: timer(False) marks a region's start; timer(True)
        waits for its end and returns the device milliseconds in between."""
        ms = ctypes.c_float(0)
        check(lib().svdag_hashset_timer(self.h, 1 if stop else 0, ctypes.byref(ms) if stop else None))
        return ms.value if stop else None

    def read(self, item_words, ptrs):
        ptrs = np.ascontiguousarray(ptrs, dtype=np.uint32)
        out = np.zeros((len(ptrs), item_words), np.uint32)
        check(lib().svdag_hashset_read(self.h, item_words, ptrs.ctypes.data_as(u32p), len(ptrs),
                                       out.ctypes.data_as(u32p)))
        return out

    def usage(self, item_words):
        out = (ctypes.c_uint32 * 4)()
        check(lib().svdag_hashset_usage(self.h, item_words, out))
        return list(out)

    def close(self):
        if self.h:
            check(lib().svdag_hashset_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def look_at(position, target, up=(0.0, 1.0, 0.0)):
    """CameraView rotation rows (row 0 = -right, row 1 = up, row 2 = forward)."""
    p = np.asarray(position, float)
    f = np.asarray(target, float) - p
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, float))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return np.concatenate([-r, u, f])


class MyGPUHashDAG:
    """Editable DAG on the GPU (reference: MyGPUHashDAG<GPU_Malloc> with
    MyGPUHashDAGUndoRedo, editMyHashDag and DAGTracer::resolve_paths)."""

    def __init__(self, levels, material_bits=4, load_factor=75, group_size=0, expected_items=0,
                 arena_slabs=0, workspace_items=0):
        cfg = Config(levels, material_bits, load_factor, group_size, expected_items, arena_slabs, workspace_items)
        self.h = ctypes.c_void_p()
        check(lib().svdag_create(ctypes.byref(cfg), ctypes.byref(self.h)))
        self.levels = levels
        self.material_bits = material_bits

    def build_voxels(self, vox):
        """Replace the DAG by a dense grid vox[x, y, z] (uint8 materials, 0xFF = empty) of side 2^k."""
        vox = np.ascontiguousarray(vox, dtype=np.uint8)
        side = vox.shape[0]
        if vox.ndim != 3 or vox.shape != (side, side, side) or side & (side - 1):
            raise ValueError("vox must be a cube of side 2^k")
        check(lib().svdag_build_voxels(self.h, vox.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(side.bit_length() - 1)))

    def build_terrain(self, seed, base, amplitude, frequency, octaves, material, band_log2=0, band_seed=0,
                      tile_log2=0):
        """band_log2 > 0 selects height-band materials (svdag_build_terrain_banded)."""
        t = Terrain(seed, base, amplitude, frequency, octaves, material)
        if band_log2:
            check(lib().svdag_build_terrain_banded(self.h, ctypes.byref(t), ctypes.c_uint32(band_log2),
                                                   ctypes.c_uint32(band_seed), ctypes.c_uint32(tile_log2)))
        elif tile_log2:
            raise ValueError("tile_log2 needs band_log2 (svdag_build_terrain tiles by itself)")
        else:
            check(lib().svdag_build_terrain(self.h, ctypes.byref(t)))

    def terrain_heights(self, x0, z0, w, d):
        out = np.empty((d, w), np.int32)
        check(lib().svdag_terrain_heights(self.h, x0, z0, w, d, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    @property
    def root(self):
        r = ctypes.c_uint32()
        check(lib().svdag_root(self.h, ctypes.byref(r)))
        return r.value

    def edit(self, kind, material, center, radius, src=(0, 0, 0), dst=(0, 0, 0), rotation=0):
        d = EditDesc(kind, material, (ctypes.c_float * 3)(*center), float(radius),
                     (ctypes.c_float * 3)(*src), (ctypes.c_float * 3)(*dst), int(rotation))
        st = EditStats()
        check(lib().svdag_edit(self.h, ctypes.byref(d), ctypes.byref(st)))
        return st

    @staticmethod
    def _descs(edits):
        # Packed with struct (a third of the time of building each field as a
        # ctypes object); kind and material wrap to uint32 as ctypes would.
        n = len(edits)
        buf = bytearray(_EDIT_PACK.size * n)
        for i, e in enumerate(edits):
            c = e[2]
            src = e[4] if len(e) > 4 else (0.0, 0.0, 0.0)
            dst = e[5] if len(e) > 5 else (0.0, 0.0, 0.0)
            rot = int(e[6]) & 0xFFFFFFFF if len(e) > 6 else 0
            _EDIT_PACK.pack_into(buf, i * _EDIT_PACK.size, int(e[0]) & 0xFFFFFFFF, int(e[1]) & 0xFFFFFFFF,
                                 c[0], c[1], c[2], e[3], src[0], src[1], src[2], dst[0], dst[1], dst[2], rot)
        return (EditDesc * n).from_buffer(buf)

    @staticmethod
    def _camera(position, rotation, fov, bmin, bmax):
        return Camera((ctypes.c_double * 3)(*position), (ctypes.c_double * 9)(*rotation), float(fov),
                      (ctypes.c_double * 3)(*bmin), (ctypes.c_double * 3)(*bmax))

    def edit_batch(self, edits):
        """edits: list of (kind, material, center, radius[, src, dst[, rotation]]); returns (stats list, gpu ms)."""
        n = len(edits)
        descs = self._descs(edits)
        stats = (EditStats * n)()
        ms = ctypes.c_float(0)
        check(lib().svdag_edit_batch(self.h, descs, n, stats, ctypes.byref(ms)))
        return list(stats), ms.value

    def edit_stream(self, edits, batch=16, collect=True):
        """edits in batches of `batch`, each followed by a collection when collect,
        enqueued back to back (svdag_edit_stream); returns (stats list, gpu ms)."""
        n = len(edits)
        descs = self._descs(edits)
        stats = (EditStats * n)()
        ms = ctypes.c_float(0)
        check(lib().svdag_edit_stream(self.h, descs, n, batch, 1 if collect else 0, stats, ctypes.byref(ms)))
        return list(stats), ms.value

    def autotune(self, sample):
        """Pick lanes per item for the merge kernels from a sample batch; returns (G, {G: ms})."""
        r = TuneResult()
        check(lib().svdag_autotune(self.h, self._descs(sample), len(sample), ctypes.byref(r)))
        return r.group_size, dict(zip((4, 8, 16, 32), list(r.ms)))

    def autotune_all(self, sample):
        """Pick lanes per item, then the per-level grid cap, from a sample
        batch; returns {"lanes_per_item": G, "max_blocks": B, "ms_by_lanes": {...},
        "ms_by_max_blocks": {...}}."""
        r = TuneReport()
        check(lib().svdag_autotune_all(self.h, self._descs(sample), len(sample), ctypes.byref(r)))
        return {"lanes_per_item": r.group_size, "max_blocks": r.max_blocks,
                "ms_by_lanes": {g: round(m, 3) for g, m in zip((4, 8, 16, 32), list(r.ms_group))},
                "ms_by_max_blocks": {b: round(m, 3) for b, m in zip((1024, 2048, 4096, 8192, 16384),
                                                                      list(r.ms_blocks))}}

    def set_load_factor(self, load_factor):
        """MyGPUHashDAG::setLoadFactor: rebuild the tables so the DAG's items fill
        load_factor percent of the head lines (handles change; history cleared)."""
        check(lib().svdag_set_load_factor(self.h, ctypes.c_uint32(int(load_factor))))

    def autotune_load_factor(self, sample, candidates=(50, 75, 95)):
        """Rebuild at each candidate load factor, time the sample batch, keep
        the fastest; returns {"load_factor": chosen, "ms_by_load_factor": {...}}."""
        c = (ctypes.c_uint32 * len(candidates))(*[int(v) for v in candidates])
        ms = (ctypes.c_float * len(candidates))()
        chosen = ctypes.c_uint32(0)
        check(lib().svdag_autotune_load_factor(self.h, self._descs(sample), len(sample), c, len(candidates),
                                               ctypes.byref(chosen), ms))
        return {"load_factor": chosen.value,
                "ms_by_load_factor": {int(v): round(m, 3) for v, m in zip(candidates, list(ms))}}

    def set_batch_mode(self, mode):
        """0: edits of a batch one after another; 1: fused waves (default);
        2: fused waves replayed from captured HIP graphs."""
        check(lib().svdag_set_batch_mode(self.h, int(mode)))

    def add_frame(self, root):
        """svdag_add_frame: the root on the undo stack (not twice in a row); clears redo."""
        check(lib().svdag_add_frame(self.h, ctypes.c_uint32(root)))

    def undo(self):
        check(lib().svdag_undo(self.h))

    def redo(self):
        check(lib().svdag_redo(self.h))

    def history(self):
        out = (ctypes.c_uint32 * 2)()
        check(lib().svdag_history(self.h, out))
        return tuple(out)

    def garbage_collect(self, mark=False, wait=True):
        """Free what the root and the fully-filled chain do not reach.
        mark=True forces the marking collection (svdag_gc_mark).  wait=False
        only enqueues the collection (svdag_gc without stats) and returns
        None: later work on the context runs after it."""
        if not wait and not mark:
            check(lib().svdag_gc(self.h, None))
            return None
        st = GcStats()
        check((lib().svdag_gc_mark if mark else lib().svdag_gc)(self.h, ctypes.byref(st)))
        return st

    def query(self, xyz, noalloc=False):
        """Material per point (NOT_FOUND where empty).  noalloc=True stages
        through the library's static per-GPU array (svdag_query_voxels_noalloc,
        at most QUERY_NOALLOC_MAX points)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32).reshape(-1, 3)
        out = np.empty(len(xyz), np.uint32)
        fn = lib().svdag_query_voxels_noalloc if noalloc else lib().svdag_query_voxels
        check(fn(self.h, xyz.ctypes.data_as(u32p), len(xyz), out.ctypes.data_as(u32p)))
        return out

    def memory_usage(self):
        m = Memory()
        check(lib().svdag_memory_usage(self.h, ctypes.byref(m)))
        return m

    def serve_replay(self, edits, arrivals, max_batch=0, frame=(0, 0), cameras=(), frame_period_us=0,
                     allow_gc=True, reserve_fraction=0.0):
        """Open-loop serving run: edits[i] arrives at arrivals[i] seconds.
        cameras: list of (position, rotation, fov, bmin, bmax) cycled per frame.
        Returns (REQUEST_DTYPE array, ServerStats)."""
        n = len(edits)
        descs = self._descs(edits)
        cams = (Camera * max(len(cameras), 1))()
        for i, c in enumerate(cameras):
            cams[i] = self._camera(*c)
        cfg = ServerConfig(int(max_batch), int(frame[0]), int(frame[1]), int(frame_period_us),
                           ctypes.cast(cams, ctypes.POINTER(Camera)), len(cameras), int(bool(allow_gc)),
                           float(reserve_fraction))
        arr = np.ascontiguousarray(arrivals, np.float64)
        assert arr.shape == (n,)
        out = np.zeros(n, REQUEST_DTYPE)
        st = ServerStats()
        check(lib().svdag_serve_replay(self.h, ctypes.byref(cfg), descs,
                                       arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                       out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
        return out, st

    def last_frame(self, w, h):
        """Hits and root of the last frame of the last serving run."""
        out = np.empty(w * h, HIT_DTYPE)
        root = ctypes.c_uint32(0)
        check(lib().svdag_server_last_frame(self.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(root)))
        return out, root.value

    def _last_pixels(self):
        return getattr(self, "_trace_pixels", 0)

    def trace(self, position, rotation, fov, bmin, bmax, w, h):
        self._trace_pixels = w * h
        cam = self._camera(position, rotation, fov, bmin, bmax)
        out = np.empty(w * h, HIT_DTYPE)
        ms = ctypes.c_float(0)
        check(lib().svdag_trace_primary(self.h, ctypes.byref(cam), w, h, out.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(ms)))
        return out, ms.value

    def trace_shadows(self, shadow_bias=1.0):
        """Sun visibility * max(N.L, 0) per pixel of the last trace(), in its layout."""
        out = np.empty(self._last_pixels(), np.float32)
        ms = ctypes.c_float(0)
        check(lib().svdag_trace_shadows(self.h, ctypes.c_float(shadow_bias), out.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(ms)))
        return out, ms.value

    def trace_ambient_occlusion(self, num_samples=3, shadow_bias=1.0, ao_ray_length=32.0, seed=1):
        """Unoccluded fraction of num_samples cosine hemisphere rays per pixel of the last trace()."""
        out = np.empty(self._last_pixels(), np.float32)
        ms = ctypes.c_float(0)
        check(lib().svdag_trace_ambient_occlusion(self.h, ctypes.c_uint32(num_samples), ctypes.c_float(shadow_bias),
                                                  ctypes.c_float(ao_ray_length), ctypes.c_uint64(seed),
                                                  out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms)))
        return out, ms.value

    def blur_ambient_occlusion(self):
        """The reference's 9-tap Gaussian over the last occlusion result, in place; returns it."""
        out = np.empty(self._last_pixels(), np.float32)
        ms = ctypes.c_float(0)
        check(lib().svdag_blur_ambient_occlusion(self.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms)))
        return out, ms.value

    def trace_colors(self, material_rgb):
        """Diffuse RGBA8 per pixel of the last trace; material_rgb: (num_materials, 3) floats."""
        pal = np.ascontiguousarray(material_rgb, np.float32).reshape(-1, 3)
        out = np.empty(self._last_pixels(), np.uint32)
        ms = ctypes.c_float(0)
        check(lib().svdag_trace_colors(self.h, pal.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(len(pal)),
                                       out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ms)))
        return out, ms.value

    def trace_lighting(self, apply_ambient_occlusion=True, apply_shadows=True, fog_density=0.0, sun_brightness=0.0):
        """Final RGBA8 per pixel of the last trace (Tracer::trace_lighting)."""
        p = Lighting(int(bool(apply_ambient_occlusion)), int(bool(apply_shadows)), float(fog_density),
                     float(sun_brightness))
        out = np.empty(self._last_pixels(), np.uint32)
        ms = ctypes.c_float(0)
        check(lib().svdag_trace_lighting(self.h, ctypes.byref(p), out.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.byref(ms)))
        return out, ms.value

    def trace_path_tracing(self, material_rgb, env=1.0, samples=1, depth=2, integrate=False, seed=1, accumulate=True):
        """Tracer::trace_path_tracing over the last trace; returns (RGBA8 image, samples accumulated, ms)."""
        pal = np.ascontiguousarray(material_rgb, np.float32).reshape(-1, 3)
        prm = PathTracing(float(env), int(samples), int(depth), int(bool(integrate)), int(seed), int(bool(accumulate)), 0)
        out = np.empty(self._last_pixels(), np.uint32)
        acc = ctypes.c_uint32(0)
        ms = ctypes.c_float(0)
        check(lib().svdag_trace_path_tracing(self.h, ctypes.byref(prm), pal.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_uint32(len(pal)), out.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.byref(acc), ctypes.byref(ms)))
        return out, acc.value, ms.value

    def canonical(self, dump=False):
        hsh = ctypes.c_uint64()
        n = ctypes.c_size_t(0)
        counts = (ctypes.c_uint32 * 32)()
        check(lib().svdag_export_canonical(self.h, ctypes.byref(hsh), None, ctypes.byref(n), counts))
        if not dump:
            return hsh.value, list(counts)
        buf = np.empty(n.value, np.uint32)
        check(lib().svdag_export_canonical(self.h, ctypes.byref(hsh), buf.ctypes.data_as(u32p), ctypes.byref(n),
                                           counts))
        return hsh.value, list(counts), buf

    def import_canonical(self, dump):
        d = np.ascontiguousarray(dump, np.uint32)
        check(lib().svdag_import_canonical(self.h, d.ctypes.data_as(u32p), ctypes.c_size_t(d.size)))

    def save(self, path):
        check(lib().svdag_save(self.h, os.fsencode(path)))

    def load(self, path):
        check(lib().svdag_load(self.h, os.fsencode(path)))

    def write_reference_dag(self):
        """The reference's MyGPUHashDAG::writeTo stream (version 3) of the reachable DAG, as bytes."""
        n = ctypes.c_size_t(0)
        check(lib().svdag_write_reference_dag(self.h, None, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        check(lib().svdag_write_reference_dag(self.h, buf.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
        return buf[:n.value].tobytes()

    def read_reference_dag(self, data):
        """MyGPUHashDAG::readFrom: imports one version-3 stream from the start of data (bytes);
        returns the number of bytes it took."""
        b = np.frombuffer(data, np.uint8)
        used = ctypes.c_size_t(0)
        check(lib().svdag_read_reference_dag(self.h, b.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(b.size),
                                             ctypes.byref(used)))
        return used.value

    def write_reference_scene(self, path, bounds=None, view=None):
        """Engine::writeTo: DAGInfo bounds (6 doubles), the DAG, CameraView (12 doubles), no textures."""
        b = None if bounds is None else (ctypes.c_double * 6)(*bounds)
        v = None if view is None else (ctypes.c_double * 12)(*view)
        check(lib().svdag_write_reference_scene(self.h, os.fsencode(path), b, v))

    def read_reference_scene(self, path):
        """Engine::readFrom: imports the scene file's DAG; returns (bounds, view)."""
        b = (ctypes.c_double * 6)()
        v = (ctypes.c_double * 12)()
        check(lib().svdag_read_reference_scene(self.h, os.fsencode(path), b, v))
        return tuple(b), tuple(v)

    def export_records(self):
        n = ctypes.c_size_t(0)
        check(lib().svdag_export_tables(self.h, None, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint32)
        check(lib().svdag_export_tables(self.h, buf.ctypes.data_as(u32p), ctypes.byref(n)))
        return buf

    def close(self):
        if self.h:
            check(lib().svdag_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
