/*
 * svdag.h - C ABI of the MI355X (gfx950) SVDAG editing back-end.
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (SVDAG_OK = 0, negative = error; svdag_last_error() gives the message).
 * Buffers are host memory unless the parameter name ends in _dev.
 * One context = one HIP stream; calls on one context must not overlap.
 *
 * Each entry point names the reference interface (mathijs727/GPU-SVDAG-Editing,
 * file:line) it stands in for.
 */
#ifndef SVDAG_H
#define SVDAG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SVDAG_OK = 0,
    SVDAG_E_INVALID = -1,   /* bad argument */
    SVDAG_E_OOM = -2,       /* slab arena exhausted (device flag); DAG unchanged */
    SVDAG_E_DEVICE = -3,    /* HIP runtime error */
    SVDAG_E_CAPACITY = -4,  /* edit workspace overflow; DAG unchanged */
    SVDAG_E_EMPTY = -5,     /* operation on an empty DAG */
    SVDAG_E_NO_DEVICE = -6  /* no GPU visible */
};

/* Message of the last failed call on this thread. */
const char* svdag_last_error(void);
/* ABI version (bumped on incompatible changes). */
uint32_t svdag_abi_version(void);

/* ------------------------------------------------------------------ device memory
 * Thin plumbing for callers without their own device allocator.
 * (reference: Memory::malloc / GpuMemoryPool, memory.h:29-63, 192-224)            */
int svdag_device_count(int* out_count);
int svdag_dev_alloc(size_t bytes, void** out_ptr_dev);
int svdag_dev_free(void* ptr_dev);
int svdag_memcpy_h2d(void* dst_dev, const void* src, size_t bytes);
int svdag_memcpy_d2h(void* dst, const void* src_dev, size_t bytes);
int svdag_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes);
int svdag_device_synchronize(void);

/* ------------------------------------------------------------------ memory-system probes
 * No reference counterpart: the practical ceilings the slab-hash find is
 * measured against.  SVDAG_PROBE_STREAM reads `bytes` sequentially with
 * 16-byte loads; SVDAG_PROBE_RANDOM_LINES reads `reads` random 128-byte lines
 * of a `bytes` buffer, their first 20 bytes (the tags and the link: the
 * words a find probe loads, all in the line's first 32-byte sector).
 * *ms is the best of five launches (HIP events on the probe's own stream).    */
#define SVDAG_PROBE_STREAM 0u
#define SVDAG_PROBE_RANDOM_LINES 1u
int svdag_hbm_probe(uint32_t kind, uint64_t bytes, uint64_t reads, float* ms);

/* ------------------------------------------------------------------ slab hash tables
 * A set of slab hash tables, one per item size in [min_words, max_words].
 * Stands in for CompactAccelerationHashTable + GpuHashTableBase
 * (hash_tables/compact_acceleration_hash_table.h:13-402, gpu_hash_table_base.h:51-326)
 * and the HashMapBench bulk harness (extras/hashmap_bench/src/main.cu:124-370).      */
typedef struct svdag_hashset svdag_hashset;

/* A run of items of one size packed in a word buffer. */
typedef struct {
    uint32_t item_words;   /* words per item */
    uint32_t count;        /* number of items */
    uint32_t word_offset;  /* first word in the word buffer */
    uint32_t out_offset;   /* first result slot */
} svdag_segment;

/* Variant of a bulk operation: 0 = one lane per item (cf. add/find),
 * 4/8/16/32 = that many lanes of a wave64 per item (cf. addAsWarp/findAsWarp),
 * 1/2/3 = find only, one lane per item with blocks assigned to segments and
 * 1/2/4 finds per lane in flight (at most 32 segments),
 * N*100 + G (204, 404, 208, 408) = find only, G lanes per group with N
 * needles per group in flight,
 * 1004/1008/1204/1208 = find only, needles binned by XCD first. */
enum { SVDAG_LANE = 0 };

/* reference: CompactAccelerationHashTable::allocate(numBuckets, numElements, itemSize)
 * (compact_acceleration_hash_table.cu:5-13) for each size.                           */
int svdag_hashset_create(uint32_t min_words, uint32_t max_words, const uint32_t* buckets_per_size,
                         const uint32_t* reserve_items_per_size, svdag_hashset** out);
int svdag_hashset_destroy(svdag_hashset* hs);
/* Bulk insert (multiset semantics); out_ptrs_dev[i] = table pointer of item i.
 * reference: add / addAsWarp (compact_acceleration_hash_table.h:27-121).
 * elapsed_ms (optional) receives the kernel time measured with HIP events.           */
int svdag_hashset_add_dev(svdag_hashset* hs, const uint32_t* words_dev, const svdag_segment* segs, uint32_t nseg,
                          uint32_t* out_ptrs_dev, int variant, float* elapsed_ms);
/* Bulk find; out_ptrs_dev[i] = pointer of a stored equal item or 0xFFFFFFFF.
 * reference: find / findAsWarp (compact_acceleration_hash_table.h:205-280).
 * elapsed_ms NULL (variants below 1000): the call returns once the find is
 * queued on the set's stream, as the reference's kernel launch does; later
 * calls on the set run after it, and svdag_hashset_timer or
 * svdag_device_synchronize waits for it.  Otherwise the call waits.          */
int svdag_hashset_find_dev(svdag_hashset* hs, const uint32_t* words_dev, const svdag_segment* segs, uint32_t nseg,
                           uint32_t* out_ptrs_dev, int variant, float* elapsed_ms);
/* Device time of a region of calls on the set's stream, by HIP events (no
 * reference counterpart): stop = 0 marks the start; stop = 1 marks the end,
 * waits for it and returns the milliseconds in between.                      */
int svdag_hashset_timer(svdag_hashset* hs, uint32_t stop, float* elapsed_ms);
/* Decode pointers to item words (host out buffer n*item_words).
 * reference: decodePointer (compact_acceleration_hash_table.h:398-402).              */
int svdag_hashset_read(svdag_hashset* hs, uint32_t item_words, const uint32_t* ptrs, uint32_t n, uint32_t* out_words);
/* Slab usage of one table: out[0]=buckets, out[1]=slab capacity, out[2]=slabs ever
 * handed out (head + overflow, recycled ones counted once), out[3]=oom flag.  reference: memory_used_by_slabs (gpu_hash_table_base.cu:187-206). */
int svdag_hashset_usage(svdag_hashset* hs, uint32_t item_words, uint32_t out[4]);

/* ------------------------------------------------------------------ editable DAG
 * reference: EDag::MyGpuDag = MyGPUHashDAG<EMemoryType::GPU_Malloc>
 * (dags/my_gpu_dags/my_gpu_hash_dag/my_gpu_hash_dag.h:59-339; engine.h:31-37).       */
typedef struct svdag_ctx svdag_ctx;

typedef struct {
    uint32_t levels;           /* D: the voxel grid is 2^D per side (reference MAX_LEVELS) */
    uint32_t material_bits;    /* 0, 4 or 8 (reference HASH_DAG_MATERIAL_BITS) */
    uint32_t load_factor;      /* percent of each bucket's head-line slots filled at expected_items (0 = 75;
                                  replaces the reference's TARGET_LOAD_FACTOR=96 items per 32-slot bucket) */
    uint32_t group_size;       /* lanes per item in DAG find/insert: 0 = default, 1, 4, 8, 16, 32 */
    uint64_t expected_items;   /* expected DAG items, sizes the bucket count (0 = 1<<20) */
    uint64_t arena_slabs;      /* overflow slab reserve per table (0 = buckets/2 + 4096) */
    uint64_t workspace_items;  /* edit workspace capacity in work items (0 = 1<<22) */
} svdag_config;

/* reference: MyGPUHashDAG::allocate (my_gpu_hash_dag.cu:24-66) incl. the fully-filled chain. */
int svdag_create(const svdag_config* cfg, svdag_ctx** out);
/* reference: MyGPUHashDAG::free (my_gpu_hash_dag.cu:68-76). */
int svdag_destroy(svdag_ctx* ctx);

/* Seeded fBm heightfield terrain (no reference counterpart: the reference
 * imports scene files through MyGPUHashDAGFactory::load_from_DAG,
 * my_gpu_hash_dag_factory.cpp:426-508, whose fully-filled rule (357-360,
 * 406-407) the built DAG follows).  Voxel (x,y,z) is solid iff y <= h(x,z),
 * h = floor(base + amplitude * fbm(x*freq, z*freq)); material = material.     */
typedef struct {
    uint64_t seed;
    double base;
    double amplitude;
    double frequency;          /* cycles per voxel of octave 0 */
    uint32_t octaves;
    uint32_t material;
} svdag_terrain;
int svdag_build_terrain(svdag_ctx* ctx, const svdag_terrain* terrain);
/* The same terrain with height-band materials (no reference counterpart: a
 * synthetic scene with one material per cube of 2^band_log2 voxels, for the
 * 8-bit-material depth-17 workload).  Voxel (x,y,z) below the surface holds
 *   (material + (y >> b) + (mix32(band_seed ^ (x>>b)*0x9E3779B1 ^ (z>>b)*0x85EBCA77) & 3))
 *   & (2^material_bits - 1),   b = band_log2 in [2, levels],
 * mix32(h) = ((h ^ h>>15) * 0x2C1B3C6D) ^ (that >> 12).  The build runs one edit
 * per column tile of 2^tile_log2 voxels (0 = levels-2 above depth 15, else one
 * edit), which keeps one edit's work items within the workspace.  The tile
 * widens by itself to the widest fully solid node (read from the height
 * pyramid), so the DAG always equals the single-edit build's.
 * svdag_build_terrain tiles the same way.                                   */
int svdag_build_terrain_banded(svdag_ctx* ctx, const svdag_terrain* terrain, uint32_t band_log2, uint32_t band_seed,
                               uint32_t tile_log2);
/* Heights of a rectangle [x0,x0+w) x [z0,z0+d), row-major in z (out: w*d int32). */
int svdag_terrain_heights(svdag_ctx* ctx, uint32_t x0, uint32_t z0, uint32_t w, uint32_t d, int32_t* out);

/* Build the DAG from a dense voxel grid of side 2^side_log2 (<= levels and
 * <= 11) at the origin: materials[(x * side + y) * side + z] is the voxel's
 * material, 0xFF = empty.  The grid replaces the DAG (history cleared); a
 * material the material bits cannot hold is SVDAG_E_INVALID.  Every node
 * whose voxels all hold one material is that material's fully-filled node,
 * the rule of the reference's factory.
 * reference: MyGPUHashDAGFactory::load_from_DAG / create_dag_cpu
 * (my_gpu_hash_dag_factory.cpp:317-508), here from a voxel source.          */
int svdag_build_voxels(svdag_ctx* ctx, const uint8_t* materials, uint32_t side_log2);

/* reference: MyGPUHashDAG::get_first_node_index / firstNodeIndex (my_gpu_hash_dag.h:72, 290-293). */
int svdag_root(svdag_ctx* ctx, uint32_t* out_root);

/* Edit tools.  reference: my_gpu_dag_editors.h
 *   SPHERE_ADD/REMOVE  MyGpuSphereEditor<bool>(center, radius, material)  (457-496)
 *   BOX_ADD/REMOVE     MyGpuBoxEditor<bool>(center, radius, material)     (82-119)
 *   PAINT              MyGpuSpherePaintEditor(center, radius, color, mat) (498-536)
 *   COPY               MyGpuCopyEditor(dag, src, dst, center, radius)     (127-253)  */
enum {
    SVDAG_SPHERE_ADD = 0,
    SVDAG_SPHERE_REMOVE = 1,
    SVDAG_BOX_ADD = 2,
    SVDAG_BOX_REMOVE = 3,
    SVDAG_PAINT = 4,
    SVDAG_COPY = 5
};
typedef struct {
    uint32_t kind;
    uint32_t material;
    float center[3];
    float radius;
    float src[3];   /* COPY only */
    float dst[3];   /* COPY only */
    /* COPY only: quarter turns (0..3) about +Y applied to the pasted region.
     * 0 is the reference's translating copy, bit for bit.  No reference
     * counterpart for 1..3 (MyGpuCopyEditor only translates): with
     * c = dst + (center - src), pivot p = floor(c + 0.5) and
     * q = p + trunc(src - dst), every destination voxel v whose unit box
     * overlaps [c - radius, c + radius) takes source voxel
     * max(q + R^k(v - p), 0), R(x, y, z) = (-z - 1, y, x), when that voxel is
     * set, and keeps its own value otherwise.  Other kinds require 0.     */
    uint32_t rotation;
} svdag_edit_desc;

typedef struct {
    uint32_t svo_items[32];    /* edit-SVO items per level (index = level) */
    uint32_t applied;          /* 0 when the edit made no leaf (reference skips the merge) */
    uint32_t root_before;
    uint32_t root_after;
    float gpu_ms;              /* device time of the edit (HIP events) */
} svdag_edit_stats;

/* Apply one edit synchronously and record the previous root as an undo frame.
 * reference: editMyHashDag<Editor>(editor, dag, undoRedo, stats, memPool, stream)
 * (my_gpu_hash_dag_edits.h:12-13; my_gpu_hash_dag_edits.cu:633-671).  stats may be NULL. */
int svdag_edit(svdag_ctx* ctx, const svdag_edit_desc* edit, svdag_edit_stats* stats);
/* svdag_edit ordered after the work already queued on the caller's HIP
 * stream (a hipStream_t passed as void*; NULL = no ordering), the stream
 * argument of editMyHashDag (my_gpu_hash_dag_edits.h:12-13).  The edit runs
 * on the context's own stream and is complete when the call returns, as the
 * reference's edit is (it synchronises the device several times per edit). */
int svdag_edit_on_stream(svdag_ctx* ctx, const svdag_edit_desc* edit, svdag_edit_stats* stats, void* stream);

/* Apply n edits in order with no host round trip between them (each edit's
 * kernels read the root the previous edit wrote on the device); one sync at
 * the end.  Same result and undo frames as n calls of svdag_edit.  No
 * reference counterpart (the reference syncs several times per edit,
 * my_gpu_hash_dag_edits.cu:526-626).  stats: n entries or NULL. */
int svdag_edit_batch(svdag_ctx* ctx, const svdag_edit_desc* edits, uint32_t n, svdag_edit_stats* stats, float* gpu_ms);

/* An edit stream: edits [0, n) in batches of `batch`, each batch followed by
 * a counting collection (svdag_gc) when collect != 0, enqueued back to back
 * with one wait at the end, so the host plans the next batches while the
 * device runs the first.  Same DAG as svdag_edit_batch per batch (and
 * svdag_gc after each when collect).  With collect the history is cleared
 * (every batch ends in a collection); without, each edit leaves its undo
 * frame.  An edit that cannot complete on the device (a workspace or arena
 * that must grow) stops the stream's collections there, and the rest runs
 * batch by batch.  No reference counterpart: the reference's edit loop
 * (editMyHashDag then garbageCollect, my_gpu_hash_dag_edits.cu:633-671,
 * my_gpu_base_dag.h:66-77) waits for every step.  Refused while a server
 * runs on the context.  stats: n entries or NULL; gpu_ms: the stream's
 * device time. */
int svdag_edit_stream(svdag_ctx* ctx, const svdag_edit_desc* edits, uint32_t n, uint32_t batch, uint32_t collect,
                      svdag_edit_stats* stats, float* gpu_ms);

/* How svdag_edit_batch schedules a batch (no reference counterpart; the
 * result is the same in every mode):
 *   SVDAG_BATCH_SEQUENTIAL  each edit's kernels one after another;
 *   SVDAG_BATCH_FUSED       (default) runs of sphere / box / paint edits whose
 *                           regions do not collide share one edit-SVO build
 *                           and one merge, and a single-wave kernel replays
 *                           the top levels edit by edit.  Copy edits run on
 *                           their own;
 *   SVDAG_BATCH_FUSED_GRAPH the fused form with each wave's launch sequence
 *                           captured once as a HIP graph (per split level and
 *                           lanes per item) and replayed. */
enum { SVDAG_BATCH_SEQUENTIAL = 0, SVDAG_BATCH_FUSED = 1, SVDAG_BATCH_FUSED_GRAPH = 2 };
int svdag_set_batch_mode(svdag_ctx* ctx, uint32_t mode);

/* Autotuner (no reference counterpart; the reference fixes warp-cooperative
 * probing at compile time).  Times a sample batch under each lanes-per-item
 * choice for the merge kernels (4, 8, 16, 32) and keeps the fastest for the
 * context.  The sample runs once untimed first; after every run the root is
 * restored and a garbage collection frees what the sample inserted, so each
 * timed run inserts its items as a real edit would and the DAG's contents
 * do not change.  Like svdag_gc, this clears the undo / redo history. */
typedef struct {
    uint32_t group_size;       /* chosen lanes per item */
    float ms[4];               /* best device ms of the sample for 4, 8, 16, 32 */
} svdag_tune_result;
int svdag_autotune(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, svdag_tune_result* out);

/* The same tuner over two knobs in turn (no reference counterpart; the
 * reference fixes both at compile time): lanes per item of the merge kernels,
 * then the grid cap of the per-level launches (the most workgroups one level's
 * kernel is launched with: 1024, 2048, 4096, 8192 or 16384; its grid-stride
 * loops cover the rest).  Each value is timed as svdag_autotune times one. */
typedef struct {
    uint32_t group_size;       /* chosen lanes per item */
    uint32_t max_blocks;       /* chosen grid cap */
    float ms_group[4];         /* best device ms of the sample for 4, 8, 16, 32 lanes */
    float ms_blocks[5];        /* the same for grid caps 1024 .. 16384, with the chosen lanes */
} svdag_tune_report;
int svdag_autotune_all(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, svdag_tune_report* out);

/* MyGPUHashDAG::setLoadFactor (my_gpu_hash_dag.cu:139-155): rebuilds the item
 * tables with bucket counts sized for the DAG's own items and copies the
 * reachable DAG into them.  load_factor (5..100) is the percentage of
 * head-line slots the items fill, as svdag_config.load_factor (the reference
 * counts items per 32-slot bucket instead, TARGET_LOAD_FACTOR = 96).  The
 * canonical DAG is unchanged; handles are reissued, so the undo / redo
 * history is cleared and the next collection marks.  On failure the old
 * tables and DAG stay.  SVDAG_E_INVALID while a server runs. */
int svdag_set_load_factor(svdag_ctx* ctx, uint32_t load_factor);

/* The autotuner's bucket-count knob (no reference counterpart; the reference
 * sets TARGET_LOAD_FACTOR at compile time, gpu_hash_dag_definitions.h:4, and
 * its hash-table sweeps pick it offline): for each candidate load factor
 * (at most 8) the tables are rebuilt as svdag_set_load_factor does and the
 * sample is timed as svdag_autotune times one knob value; the fastest is
 * left set.  ms (optional, num_candidates entries) receives the best device
 * ms of each. */
int svdag_autotune_load_factor(svdag_ctx* ctx, const svdag_edit_desc* sample, uint32_t n, const uint32_t* candidates,
                               uint32_t num_candidates, uint32_t* chosen, float* ms);

/* reference: MyGPUHashDAGUndoRedo::add_frame (my_gpu_base_dag.h:23-33): root
 * goes on the undo stack unless it is the last frame there, and the redo stack
 * is cleared.  Every edit already adds its own frame (editMyHashDag does).
 * root must be a root this context held (the current one or one still on a
 * stack): a collection frees what only other roots reach. */
int svdag_add_frame(svdag_ctx* ctx, uint32_t root);
/* reference: MyGPUHashDAGUndoRedo::undo / redo (my_gpu_base_dag.h:35-64). */
int svdag_undo(svdag_ctx* ctx);
int svdag_redo(svdag_ctx* ctx);
/* Undo/redo stack depths: out[0] = undo frames, out[1] = redo frames. */
int svdag_history(svdag_ctx* ctx, uint32_t out[2]);

/* reference: MyGPUHashDAGUndoRedo::garbageCollect -> MyGPUHashDAG::garbageCollect
 * (my_gpu_base_dag.h:66-77; my_gpu_hash_dag.cu:364-475).  Clears the history
 * and frees every stored item that the root and the fully-filled chain do
 * not reach (the reference marks and, only under CAPTURE_MEMORY_STATS_SLOW,
 * sweeps; this always frees).  The same items are freed either way, but the
 * work differs: every stored item keeps a count of the references other
 * stored items hold to it, so the collection releases the roots held since
 * the last one and frees what drops to zero, in time proportional to what
 * died.  After an edit was rolled back or a dump was imported the counts are
 * rebuilt: that collection marks from the roots, sweeps and recounts, like
 * svdag_gc_mark.  A count that reached its ceiling (64) stays there, so its
 * item is freed only by a marking collection; svdag_gc marks instead once
 * the stored items exceed 1.5 times the live count of the last marking
 * collection.  items_live counts the stored items after the collection
 * (all reachable after a marking one); marked = 1 if it marked.
 * stats == NULL: a counting collection is only enqueued on the context's
 * stream and the call returns without waiting for it (everything later on
 * the context runs after it); the next collection takes in its results, and
 * if it stopped short (a list overflowed) that next collection marks.  A
 * marking collection always completes before the call returns. */
typedef struct {
    uint64_t items_freed;
    uint64_t slabs_freed;
    uint64_t items_live;
    float gpu_ms;
    uint32_t marked;
} svdag_gc_stats;
int svdag_gc(svdag_ctx* ctx, svdag_gc_stats* stats);
/* The marking collection itself, whatever the counts say: marks from the
 * roots level by level (the reference's scheme, my_gpu_hash_dag.cu:377-463),
 * sweeps every line and recounts the references.  No reference counterpart
 * as a separate call; it checks svdag_gc (a marking collection right after a
 * counting one frees nothing). */
int svdag_gc_mark(svdag_ctx* ctx, svdag_gc_stats* stats);

/* Voxel queries: xyz = n*3 uint32 (host); out_material[i] = material or 0xFFFFFFFF.
 * reference: DAGUtils::get_value (dag_utils.h:146-182) / DAGTracer::get_voxel_values
 * (dag_tracer.cu:361-390). */
int svdag_query_voxels(svdag_ctx* ctx, const uint32_t* xyz, uint32_t n, uint32_t* out_material);

/* svdag_query_voxels without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results; the points and results are staged through one
 * static device array per GPU, set up at svdag_create, so n is at most
 * 1048576 (larger n is rejected with SVDAG_E_INVALID).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (from different contexts
 * or threads) race. */
int svdag_query_voxels_noalloc(svdag_ctx* ctx, const uint32_t* xyz, uint32_t n, uint32_t* out_material);

/* Primary rays.  reference: DAGTracer::resolve_paths -> Tracer::trace_paths
 * (dag_tracer.cu:59-134; tracer.cu:186-273, 353-411).  Camera as CameraView
 * (camera_view.h): position, 3x3 rotation (row 0 = -right, row 1 = up,
 * row 2 = forward), fov in degrees; bounds as DAGInfo (dag_info.h). */
typedef struct {
    double position[3];
    double rotation[9];
    double fov_degrees;
    double bounds_min[3];
    double bounds_max[3];
} svdag_camera;
/* One hit record per pixel, stored at row (height-1-y) like the reference. */
typedef struct {
    uint32_t x, y, z;          /* hit voxel path (0,0,0 on a miss) */
    uint32_t material;         /* material id, 0xFFFFFFFF on a miss */
    float t;                   /* ray parameter of the entry face */
    float px, py, pz;          /* hit position */
} svdag_hit;
int svdag_trace_primary(svdag_ctx* ctx, const svdag_camera* camera, uint32_t width, uint32_t height,
                        svdag_hit* out_hits, float* gpu_ms);

/* Shadows and ambient occlusion of the surface the last svdag_trace_primary
 * call on this context found (same size, same root; nothing else traced in
 * between).  Output: width*height floats in the hit buffer's layout; pixels
 * without a surface (a miss, or the null path (0,0,0) as Path::is_null
 * treats it) get 0.
 * svdag_trace_shadows: sun visibility times max(N.L, 0), L = the reference's
 *   sun direction normalize(0.3, 1, 0.5); the shadow ray starts at
 *   position + shadow_bias * L.
 *   reference: DAGTracer::resolve_shadows (dag_tracer.cu:198-221) ->
 *   Tracer::trace_shadows (tracer.cu:613-633).
 * svdag_trace_ambient_occlusion: the fraction of num_samples cosine-weighted
 *   hemisphere rays (length ao_ray_length, from position + normal *
 *   shadow_bias) that reach no voxel; the samples of stored pixel i come from
 *   PCG32 stream i * max(seed, 1).
 *   reference: DAGTracer::resolve_ambient_occlusion (dag_tracer.cu:163-187) ->
 *   Tracer::trace_ambient_occlusion (tracer.cu:661-691).                      */
int svdag_trace_shadows(svdag_ctx* ctx, float shadow_bias, float* out_sun, float* gpu_ms);
int svdag_trace_ambient_occlusion(svdag_ctx* ctx, uint32_t num_samples, float shadow_bias, float ao_ray_length,
                                  uint64_t seed, float* out_ao, float* gpu_ms);

/* DAGTracer::resolve_ambient_occlusion_blur (dag_tracer.cu:189-196) ->
 * BlurKernel::apply (blur.cu:68-85) with the tracer's 9-tap kernel
 * (half size 4, dag_tracer.cu:19): a horizontal then a vertical pass over the
 * last ambient-occlusion result, in place; taps outside the image are
 * skipped.  width and height must be multiples of 8.  out (optional)
 * receives the blurred buffer in the hit buffer's layout. */
int svdag_blur_ambient_occlusion(svdag_ctx* ctx, float* out, float* gpu_ms);

/* DAGTracer::resolve_colors -> Tracer::trace_colors -> computeColorAtSurface
 * (dag_tracer.cu:137-161; tracer.cu:413-596) for a DAG with materials.  The
 * reference samples each material's top / side / all textures; texture
 * images are not part of this library, so each material is one colour:
 * material_rgb[3 m .. 3 m + 2] (linear 0..1 floats, as the texture fetch
 * returns), for m < num_materials.  Per pixel of the last primary trace: no
 * surface -> 0; a material without a colour -> the reference's invalid-colour
 * checker (1, b, 1 - b), b = (x ^ y ^ z) & 1 of the hit voxel; else the
 * material's colour.  Stored (and copied to out, if given) as
 * ColorUtils::float3_to_rgb888: R | G << 8 | B << 16 | 0xFF << 24. */
int svdag_trace_colors(svdag_ctx* ctx, const float* material_rgb, uint32_t num_materials, uint32_t* out_rgba8,
                       float* gpu_ms);

/* DAGTracer::resolve_lighting -> Tracer::trace_lighting (dag_tracer.cu:223-254;
 * tracer.cu:693-750): per pixel of the last primary trace, the sky colour
 * where nothing was hit, else (AO * 0.5 * sky + sun_visibility * sun colour *
 * sun_brightness) * Kd with Kd the colour pass's sRGB colour made linear,
 * optionally fogged (fog_density > 0), then made sRGB and stored as RGBA8.
 * AO is the last (blurred, if svdag_blur_ambient_occlusion ran) occlusion
 * result and sun_visibility the last svdag_trace_shadows result of the same
 * trace; each counts as 1 when not applied.  resolve_lighting leaves the
 * reference's sunBrightness unset, so its frames have sun_brightness 0; pass
 * 0 for the reference's image.  The tool overlay is not drawn. */
typedef struct {
    uint32_t apply_ambient_occlusion;
    uint32_t apply_shadows;
    float fog_density;       /* the reference's default is 0 (engine.h:214) */
    float sun_brightness;
} svdag_lighting;
int svdag_trace_lighting(svdag_ctx* ctx, const svdag_lighting* params, uint32_t* out_rgba8, float* gpu_ms);

/* DAGTracer::resolve_path_tracing -> Tracer::trace_path_tracing
 * (dag_tracer.cu:256-313; tracer.cu:758-842) over the camera and root of the
 * last primary trace: num_samples paths of at most max_path_depth bounces
 * per pixel, next-event estimation toward the sun, uniform hemisphere
 * continuation, sky light times environment_brightness; colours as
 * svdag_trace_colors (material_rgb, one colour per material).  Without
 * integrate_pixel each path starts on the stored primary surface with the
 * colour pass's colour (svdag_trace_colors must have run); with it the
 * camera ray is jittered in the pixel.  Sums accumulate across calls while
 * accumulate is set and the image, root, camera, depth, brightness and
 * integration are unchanged (else they restart, as the reference clears its
 * buffer); *accumulated (optional) receives the samples summed so far.
 * Once more than 256 are summed no new samples are traced (as there).
 * Output: min(sum / samples, 1) made sRGB, RGBA8 per stored pixel.  The
 * samples of stored pixel i come from PCG32 stream i * max(seed, 1); sin and
 * cos of the sample angle use this library's own reduction (as the
 * occlusion pass). */
typedef struct {
    float environment_brightness;
    uint32_t num_samples;
    uint32_t max_path_depth;
    uint32_t integrate_pixel;
    uint64_t seed;
    uint32_t accumulate;
    uint32_t reserved;
} svdag_path_tracing;
int svdag_trace_path_tracing(svdag_ctx* ctx, const svdag_path_tracing* params, const float* material_rgb,
                             uint32_t num_materials, uint32_t* out_rgba8, uint32_t* accumulated, float* gpu_ms);

/* Slab-arena usage over all tables of the DAG (no reference counterpart;
 * the reference reports memory_used_by_slabs, gpu_hash_table_base.cu:187-206).
 * min_free_fraction: the smallest free / capacity ratio over the tables. */
typedef struct {
    uint64_t slab_capacity;
    uint64_t slabs_used;
    uint64_t slabs_free;
    uint64_t bytes_reserved;
    double min_free_fraction;
} svdag_memory;
int svdag_memory_usage(svdag_ctx* ctx, svdag_memory* out);

/* ---------------------------------------------------------------- serving
 * Continuous batching (no reference counterpart; the reference applies one
 * edit per Engine::edit call, engine.h:104-163).  A server thread takes the
 * requests that are waiting, admits a FIFO prefix against slab-arena
 * headroom and applies it as one batch (svdag_edit_batch semantics: the
 * result equals applying the requests one by one in submission order), while
 * requests that arrive meanwhile queue for the next batch.  A frame thread
 * traces the latest published root on its own stream every frame period,
 * overlapping the edit batches.  When headroom runs out the server collects
 * garbage (clearing the undo history) and, if that is not enough, grows the
 * arenas.  While a server runs, other calls on the context stay legal; they
 * are serialised with the server's batches. */
typedef struct {
    uint32_t max_batch;            /* most requests in one batch (0 = 256) */
    uint32_t frame_width;          /* per-frame primary trace; 0 = no frames */
    uint32_t frame_height;
    uint32_t frame_period_us;      /* 0 = 16667 */
    const svdag_camera* cameras;   /* frame f uses cameras[f % num_cameras] (copied at start) */
    uint32_t num_cameras;
    uint32_t allow_gc;             /* collect garbage when headroom runs out */
    double reserve_fraction;       /* admission keeps this fraction of every table free (0 = 0.02) */
} svdag_server_config;

typedef struct {
    int32_t status;                /* SVDAG_OK or the error the batch failed with */
    uint32_t applied;
    uint32_t root_after;
    uint32_t batch;                /* index of the batch the request ran in */
    double submit_s;               /* server clock (seconds since start) */
    double admit_s;
    double done_s;
} svdag_request_info;

typedef struct {
    uint64_t requests;
    uint64_t batches;
    uint64_t frames;
    uint64_t frames_skipped;       /* frame due while the previous trace still ran */
    uint64_t gcs;
    uint64_t grows;
    uint64_t deferred;             /* admission deferrals (a request counts once per batch it missed) */
    double wall_s;
    double edit_busy_s;            /* host wall time spent inside batches */
    double trace_ms_avg;           /* device time per frame trace */
    uint32_t last_frame_root;
    uint32_t max_batch_seen;
    uint64_t sequential_spine_edits; /* fused-wave edits whose top levels the one-wave sequential
                                        spine built (the parallel spine declined their wave) */
    uint64_t handed_back_edits;      /* edits a fused wave handed back to run on their own */
} svdag_server_stats;

int svdag_server_start(svdag_ctx* ctx, const svdag_server_config* cfg);
int svdag_server_submit(svdag_ctx* ctx, const svdag_edit_desc* edit, uint64_t* ticket);
/* Blocks until the request is done. */
int svdag_server_wait(svdag_ctx* ctx, uint64_t ticket, svdag_request_info* out);
/* Drains the queue, stops both threads and reports. */
int svdag_server_stop(svdag_ctx* ctx, svdag_server_stats* stats);
/* Hit buffer (frame_width * frame_height records) and root of the last
 * traced frame; valid after svdag_server_stop until the next start. */
int svdag_server_last_frame(svdag_ctx* ctx, svdag_hit* out, uint32_t* root);
/* Open-loop driver: starts a server, submits edits[i] at arrival_s[i]
 * (seconds after start, non-decreasing), waits for all, stops.  out: n
 * entries.  The timing loop runs in native code so the arrival process is
 * not disturbed by the caller's runtime. */
int svdag_serve_replay(svdag_ctx* ctx, const svdag_server_config* cfg, const svdag_edit_desc* edits,
                       const double* arrival_s, uint32_t n, svdag_request_info* out, svdag_server_stats* stats);

/* Canonical form of the reachable DAG (nodes renumbered per level by sorted
 * content, children replaced by canonical ids).  hash = FNV-1a 64 of the dump.
 * If buf is non-NULL and *len_words is large enough the dump is written;
 * *len_words always receives the dump length.  out_counts (optional, 32
 * entries) receives unique items per level. */
int svdag_export_canonical(svdag_ctx* ctx, uint64_t* out_hash, uint32_t* buf, size_t* len_words,
                           uint32_t* out_counts);

/* Reachable DAG as flat records (for host-side checkers).  *len_words
 * receives the number of words needed; the dump is written if buf is large
 * enough.  Layout: [levels, material_bits, root, n_records] then per record
 * [level, handle, size, words[size]] for every (level, handle) reachable from
 * the root or from a level-0 fully-filled node (the garbage-collection roots).
 * Fails with SVDAG_E_INVALID if a reachable handle points at a free slot. */
int svdag_export_tables(svdag_ctx* ctx, uint32_t* buf, size_t* len_words);

/* Rebuild a DAG from a canonical dump (svdag_export_canonical's format): items
 * are found or inserted level by level, bottom up, and the root is set; the
 * undo/redo history is cleared.  The dump must come from a context with the
 * same depth and material bits.  Checkpoint / resume (no reference
 * counterpart for this format; the reference's MyGPUHashDAG::writeTo /
 * readFrom, my_gpu_hash_dag.cu:157-177, dump raw slab memory of its own
 * table layout). */
int svdag_import_canonical(svdag_ctx* ctx, const uint32_t* buf, size_t len_words);
/* File form: header {"SVDG", version 1, levels, material bits, dump length},
 * the canonical dump, then its FNV-1a 64 checksum.  svdag_load checks all of
 * them before importing. */
int svdag_save(svdag_ctx* ctx, const char* path);
int svdag_load(svdag_ctx* ctx, const char* path);

/* The reference's own on-disk format (version 3).
 *
 * svdag_read_reference_dag replaces MyGPUHashDAG::readFrom
 * (my_gpu_hash_dag.cu:167-177): it parses one MyGPUHashDAG stream from buf
 * (version, fullyFilledNodes, root handle, one 32-slot slab table per item
 * size with its slab allocator; gpu_hash_table_base.cu:138-162,
 * dynamic_slab_alloc.cu:181-188), walks every item reachable from the root,
 * and finds or inserts the same DAG in ctx, whose root it becomes (undo/redo
 * cleared, as after svdag_import_canonical).  *consumed (optional) receives
 * the stream's length in bytes.  The stream's depth (fullyFilledNodes width)
 * and material bits (its height, 2^M) must equal the context's; a truncated
 * stream, a header that does not describe 32-slot slabs of 10 + 32 s words,
 * or a reachable handle that addresses no stored item is SVDAG_E_INVALID.
 *
 * svdag_write_reference_dag replaces MyGPUHashDAG::writeTo
 * (my_gpu_hash_dag.cu:157-165) for a GPU_Malloc DAG: the reachable DAG and the
 * fully-filled chain in tables of about 96 items per bucket
 * (TARGET_LOAD_FACTOR), as MyGPUHashDAG::allocate and setLoadFactor lay them
 * out.  Same calling pattern as svdag_export_canonical: *len receives the
 * stream's length in bytes, and buf (if non-NULL and *len is large enough)
 * the stream.  A table the reference's handles cannot address is
 * SVDAG_E_CAPACITY. */
int svdag_read_reference_dag(svdag_ctx* ctx, const void* buf, size_t len, size_t* consumed);
int svdag_write_reference_dag(svdag_ctx* ctx, void* buf, size_t* len);

/* Engine::readFrom / Engine::writeTo (engine.cpp:1543-1557): the scene file
 * the reference loads at start-up (main.cpp:104-109): DAGInfo (AABB min and
 * max, 6 doubles), the MyGPUHashDAG stream, CameraView (position then the
 * row-major 3x3 rotation, 12 doubles) and VoxelTextures.  Reading imports the
 * DAG as svdag_read_reference_dag does and returns the bounds and view
 * (either may be NULL); the voxel textures are checked for extent and
 * skipped (colour is not part of this library).  Writing takes bounds and
 * view (NULL: [0, 2^depth]^3 and CameraView's defaults) and writes no
 * materials. */
int svdag_read_reference_scene(svdag_ctx* ctx, const char* path, double* bounds, double* view);
int svdag_write_reference_scene(svdag_ctx* ctx, const char* path, const double* bounds, const double* view);

#ifdef __cplusplus
}
#endif

#endif /* SVDAG_H */
