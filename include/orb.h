/*
 * orb.h -- C-ABI of the MI355X-native ORB (orthogonal recursive bisection)
 * build: liborb_hip.so, HIP kernels for gfx950.
 *
 * The entry points mirror the reference's service boundary: each service in
 * the reference is an mdl::BasicService whose
 *     int Service(PST pst, void *vin, int nIn, void *vout, int nOut)
 * takes Cell[nCells] and returns the number of bytes written into vout
 * (src/services/TraversePST.h:24-45, src/services/pst.h:66-81).  Here a
 * context replaces the PST leaf's LocalData (src/services/pst.h:6-41): it
 * owns the device-resident particles, the cell-to-range map and the scratch
 * buffers.  One context drives one GPU share; there are no ranks.
 *
 * Conventions
 *  - Cell and output buffers are caller-owned host memory.
 *  - Every entry point returns >= 0 on success (a byte count for services,
 *    0 otherwise) and a negative ORB_ERR_* code on failure; nothing calls
 *    exit() (the reference's CUDA_CHECK did, src/constants.h:19-23).
 *  - One context per host thread; contexts share no state.
 *  - Particles are float32, axis-major SoA (x[n], y[n], z[n]), as blitz's
 *    storage ordering (0,1) lays out particles(N,3) in
 *    src/services/init.cu:32-45.
 */
#ifndef ORB_H
#define ORB_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* struct Cell (src/cell.h:9-17): same fields, order, padding; 52 bytes. */
typedef struct orb_cell {
    int32_t id;
    int32_t nLeafCells;
    int32_t prevCutAxis;
    int32_t cutAxis;
    bool foundCut;
    float cutMarginLeft;
    float cutMarginRight;
    float lower[3];
    float upper[3];
} orb_cell;

/* struct META_PARAMS (src/constants.h:13-17).  o=0: all false; o=1:
 * GPU_COUNT; o=2: GPU_COUNT + GPU_PARTITION (src/orbit.cpp:48-66).       */
typedef struct orb_meta_params {
    bool GPU_COUNT;
    bool GPU_PARTITION;
    bool FAST_MEDIAN;
} orb_meta_params;

/* error codes */
enum {
    ORB_OK = 0,
    ORB_ERR_ARG = -1,     /* invalid argument / size                          */
    ORB_ERR_HIP = -2,     /* a HIP runtime call failed (orb_last_hip_error)   */
    ORB_ERR_NOMEM = -3,   /* device or host allocation failed                 */
    ORB_ERR_UB = -4,      /* input on which the reference is undefined: an
                             empty cell, a cell with no key >= cut (Hoare runs
                             outside the cell, partition.cpp:38-52), or a child
                             with no positive extent (cell.h:102-126)          */
    ORB_ERR_STATE = -5,   /* call out of sequence                             */
    ORB_ERR_NODEV = -6,   /* no GPU / HIP extension not usable                */
    ORB_ERR_BUSY = -7,    /* serving: not admitted (memory or slots)          */
};

/* service ids: enum pst_service (src/services/pst.h:66-81) */
enum {
    ORB_PST_SRV_STOP = 0,
    ORB_PST_SETADD,
    ORB_PST_INIT,
    ORB_PST_INITGPU,
    ORB_PST_COPYPARTICLES,
    ORB_PST_COPYCELLS,
    ORB_PST_COUNTLEFTGPU,
    ORB_PST_COUNTLEFTAXISGPU,
    ORB_PST_COUNTLEFT,
    ORB_PST_PARTITION,
    ORB_PST_PARTITIONGPU,
    ORB_PST_COUNT,
    ORB_PST_FINALIZE,
    ORB_PST_MAKEAXIS,
};

/* level schedules */
enum {
    ORB_LEVELS_REFERENCE = 0, /* l = 1 .. ceil(log2 d)-1 and the reference's
                                 heap slices (src/orbit.cpp:102-111): d = 2^k
                                 ends with d/2 particle ranges                 */
    ORB_LEVELS_FULL = 1,      /* split every cell with nLeafCells >= 2 until
                                 each has one leaf: d ranges for any d         */
};

/* comparison used by the count-left passes */
enum { ORB_CMP_LT = 0,  /* x <  cut: ServiceCountLeft (countLeft.cpp:35)       */
       ORB_CMP_LE = 1   /* x <= cut: the GPU services (countLeftGPUAxis.cu:160) */ };

/* partition used after each level */
enum { ORB_PART_HOARE = 0,      /* in-place Hoare, partition.cpp:28-60: exact order */
       ORB_PART_STABLE_LE = 1   /* keys <= cut first, stable (set semantics of
                                   partitionGPU.cu:58-231, deterministic order) */ };

/* build forms */
enum { ORB_FORM_PLAIN = 0,    /* host-driven: one count launch per bisection
                                 pass per level, counts read back each pass   */
       ORB_FORM_BATCHED = 1   /* device-driven: all cells of a level bisected
                                 together, multi-step replay, no host round
                                 trips inside the build                       */ };

#define ORB_MAX_PASSES 32

typedef struct orb_ctx orb_ctx;

/* ------------------------------------------------------------------ */
/* lifecycle                                                           */
/* ------------------------------------------------------------------ */

/* ServiceInit::Service (src/services/init.cu:27-144).  Allocates the
 * device-resident particle set for nParticles (one GPU = one rank's share)
 * and the cell-to-range map for d leaves, sets range(0) = [0, n).  With
 * generate=true fills x, y, z with the reference generator (xorshf96 with
 * the seeds of init.cu:11, i-major draw order, init.cu:47-53).  The
 * comparison and partition follow params as in orbit.cpp:48-66.         */
int orb_init(orb_ctx **ctx, int64_t nParticles, int32_t d, bool generate,
             orb_meta_params params);

/* ServiceFinalize::Service (src/services/finalize.cu:10-48). */
int orb_finalize(orb_ctx *ctx);

/* Override the comparison / partition chosen from META_PARAMS
 * (ORB_CMP_*, ORB_PART_*); -1 keeps the current value.                   */
int orb_set_semantics(orb_ctx *ctx, int32_t cmp, int32_t partition);

/* The generate=false input path (init.cu:54-59 loads a Tipsy file):
 * copies host SoA arrays of ctx's particle count to the device and resets
 * the range map to range(0) = [0, n).  A NaN coordinate is refused with
 * ORB_ERR_ARG (it compares false both ways, which the batched form's
 * histogram rounds cannot count as the reference's comparisons do);
 * infinities are accepted.                                               */
int orb_set_particles(orb_ctx *ctx, const float *x, const float *y, const float *z, int64_t n);
/* Copies the device particles (current order) back to host arrays.       */
int orb_get_particles(orb_ctx *ctx, float *x, float *y, float *z, int64_t n);
/* Resets the range map to range(0) = [0, n) (particles keep their order). */
int orb_reset_ranges(orb_ctx *ctx);
/* Device-side snapshot of the particle order, and its restore (which also
 * resets the range map): repeated builds from the same input without a
 * host round trip.  No counterpart in the reference (its particles are
 * generated once per run, init.cu:47-53).                                */
int orb_save_particles(orb_ctx *ctx);
int orb_restore_particles(orb_ctx *ctx);

/* ------------------------------------------------------------------ */
/* services: Cell[nCells] in, bytes written (>= 0) or ORB_ERR_* out     */
/* ------------------------------------------------------------------ */

/* ServiceCopyParticles (copyParticles.cu:10-60): particles already live on
 * the device; kept so the reference's service sequence runs unchanged.   */
int orb_copy_particles(orb_ctx *ctx);
/* ServiceCopyCells (copyCells.cu:10-64): builds the level's chunk table on
 * the device from the cells' ranges (no host block table).               */
int orb_copy_cells(orb_ctx *ctx, const orb_cell *cells, int32_t nCells);
/* ServiceCount (count.cpp:8-19): out[i] = end - begin of cells[i].id.    */
int orb_count(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out);
/* ServiceMakeAxis (makeAxis.cpp:9-33): gathers each cell's cut-axis
 * column into the device particlesT buffer over [begin, end) exactly (the
 * reference's inclusive blitz Range copies one element past the cell).   */
int orb_make_axis(orb_ctx *ctx, const orb_cell *cells, int32_t nCells);
/* ServiceCountLeft (countLeft.cpp:9-42): out[i] = #{x < cut} over the
 * cell's cut-axis column, found cells skipped (their out[i] untouched).  */
int orb_count_left(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out);
/* ServiceCountLeftGPUAxis (countLeftGPUAxis.cu:188-259): out[i] =
 * #{particlesT <= cut} for every cell (particlesT from orb_make_axis).   */
int orb_count_left_gpu_axis(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out);
/* ServiceCountLeftGPU (countLeftGPU.cu:80-162): out[i] = #{x_axis <= cut}
 * on the SoA arrays with each cell's own axis, every cell.               */
int orb_count_left_gpu(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out);
/* ServicePartition (partition.cpp:18-65): Hoare partition of every cell at
 * getCut() on cutAxis, moving x, y, z together; writes the children's
 * ranges [begin, i) and [i, end).  Reproduces the serial loop's exact
 * permutation (two scans + disjoint transpositions + the final swap).    */
int orb_partition(orb_ctx *ctx, const orb_cell *cells, int32_t nCells);
/* ServicePartitionGPU (partitionGPU.cu:282-600): keys <= cut first, then
 * keys > cut, each in input order; child ranges from the count at the
 * cut the partition uses.                                                */
int orb_partition_gpu(orb_ctx *ctx, const orb_cell *cells, int32_t nCells);

/* No counterpart in the reference: orb_count_left (cmp ORB_CMP_LT: #{x < cut},
 * found cells skipped, their out[i] untouched) or orb_count_left_gpu
 * (ORB_CMP_LE: #{x <= cut} for every cell) for callers that cannot
 * allocate (code under stream capture, tight serving loops).  It allocates
 * nothing: every argument is caller-owned DEVICE memory -- x, y, z [n]
 * (axis-major SoA), the cell-to-range map ranges [cap][2], cells [nCells],
 * out [nCells] -- and the work is enqueued on `stream` (a hipStream_t;
 * NULL: the default stream) without synchronising; out is ready when the
 * stream reaches this point.  The tables live in one static device array
 * per GPU (7.2 MB, held until the process ends), so at most 8096 cells
 * (MAX_CELLS, src/constants.h:11) and 2^30 particles; larger sizes return
 * ORB_ERR_ARG.  Calls that overlap on one GPU share the array and race.
 * Cell ranges are clamped into [0, n) and an id outside [0, cap) is an
 * empty cell; cells whose ranges overlap so much that their chunks exceed
 * the table get UINT32_MAX counts.  Returns nCells * 4.                  */
int orb_count_left_static(const float *x, const float *y, const float *z, int64_t n,
                          const uint32_t *ranges, int64_t cap, const orb_cell *cells,
                          int32_t nCells, int32_t cmp, uint32_t *out, void *stream);

/* mdl->RunService(id, nIn, vin, vout) for the services above (the ids of
 * src/services/pst.h:66-81); vin = Cell[nIn/52] except for COPYPARTICLES
 * and FINALIZE (ignored).  Returns the service's byte count.             */
int orb_run_service(orb_ctx *ctx, int32_t service_id, const void *vin, int32_t nIn,
                    void *vout, int32_t nOut);

/* ------------------------------------------------------------------ */
/* whole build (the level loop of src/orbit.cpp:74-275)                 */
/* ------------------------------------------------------------------ */

typedef struct orb_build_info {
    int32_t nlevels;          /* levels processed                          */
    int32_t total_passes;     /* bisection passes summed over levels       */
    int32_t level_passes[64]; /* passes per level (<= 32 each)             */
    int32_t error_cell;       /* heap id that raised ORB_ERR_UB, else -1: the first
                                 undefined cell in the reference's order (the lowest
                                 level; there a cell whose split fails before one
                                 whose partition does, orbit.cpp:235-274; then the
                                 lowest heap id), in both build forms              */
    int32_t pad;
    double ms_total;          /* host wall time of the build               */
} orb_build_info;

/* Number of heap slots for d leaves: max(2d-1, 2^(ceil(log2 d)+1)-1).    */
int64_t orb_heap_capacity(int32_t d);
/* Runs the whole build on the device-resident particles, starting from
 * range(0) = [0, n) and the root cell [-0.5, 0.5]^3 with cutAxis 0.       */
int orb_build(orb_ctx *ctx, int32_t form, int32_t levels_mode, orb_build_info *info);
/* Results of the last build: the cell heap (by heap id, unused slots
 * zero), the cell-to-range map ([cap][2]), and which ids exist.  After a
 * batched build the tree stays in HBM; these getters copy it back.       */
int orb_get_heap(orb_ctx *ctx, orb_cell *heap, int64_t cap);
int orb_get_ranges(orb_ctx *ctx, uint32_t *ranges, int64_t cap);
int orb_get_exists(orb_ctx *ctx, uint8_t *exists, int64_t cap);
/* Trace of per-pass left counts: [cap][32], UINT32_MAX where a cell was
 * not counted.  Enable before orb_build; costs host memory and copies.   */
int orb_enable_trace(orb_ctx *ctx, bool on);
int orb_get_pass_counts(orb_ctx *ctx, uint32_t *counts, int64_t cap);

/* ------------------------------------------------------------------ */
/* tree export.  No counterpart in the reference: it never outputs the   */
/* tree (src/orbit.cpp:284-286).  A leaf is an existing cell whose        */
/* children do not exist.                                               */
/* ------------------------------------------------------------------ */

/* Particle ids: when on, an id array (the input index of each particle)
 * moves with the particles through every partition; it restarts as the
 * identity at orb_set_particles and when turned on, and is saved and
 * restored with the particles.  Off by default (it adds 4 B per moved
 * particle).                                                            */
int orb_enable_ids(orb_ctx *ctx, bool on);
int orb_get_ids(orb_ctx *ctx, uint32_t *ids, int64_t n);
/* Leaves of the last build in heap-id order: their ids and [begin, end)
 * ranges ([cap][2]).  *count gets the number of leaves; ORB_ERR_ARG if
 * cap is smaller (ids/ranges may be NULL to ask for the count).          */
int orb_get_leaves(orb_ctx *ctx, int32_t *ids, uint32_t *ranges, int64_t cap, int64_t *count);
/* Leaf heap id of every particle (-1 outside every leaf), indexed by its
 * current position, or by its input index when by_input_order (needs
 * orb_enable_ids before the build).                                     */
int orb_get_membership(orb_ctx *ctx, int32_t *leaf_of, int64_t n, bool by_input_order);
/* Tight bounding box of each leaf's particles ([count][3] each, leaves in
 * orb_get_leaves order): segmented min/max on the device.  An empty leaf
 * gets lower = +inf, upper = -inf.                                      */
int orb_get_leaf_bounds(orb_ctx *ctx, float *lower, float *upper, int64_t cap);

/* ------------------------------------------------------------------ */
/* workload weights.  The thesis defines the weight w(p_i) of a particle */
/* as its share of the work (the simulation steps it needs) and the      */
/* split target as a weight quantile (documentation/notes.tex:259-271,   */
/* 326-336); the reference's code counts particles only.  With weights   */
/* on, both build forms run the reference's controller on weight sums:   */
/* count(x < cut) becomes the weight of those keys and the cell's count  */
/* its weight (uint64 sums), the termination rule of orbit.cpp:204-208   */
/* is evaluated in the same float32 roundings (the uint64 sums rounded   */
/* to nearest) and |difference| < 3 weight units ends the search.  With  */
/* every weight 1 a weighted build equals the unweighted one.  Weights   */
/* move with the particles through every partition.                      */
/* ------------------------------------------------------------------ */
/* Sets the weights (uint32, in the particles' current order; n = the
 * context's particle count); NULL turns weights off.  orb_set_particles
 * turns them off (new particles need their own).  orb_save_particles
 * saves them with the particles; orb_restore_particles then restores them
 * (ORB_ERR_STATE if weights were set after the snapshot).                */
int orb_set_weights(orb_ctx *ctx, const uint32_t *w, int64_t n);
int orb_get_weights(orb_ctx *ctx, uint32_t *w, int64_t n);
/* ServiceCount with weights: out[i] = the weight of cells[i]'s particles. */
int orb_count_weight(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint64_t *out);
/* ServiceCountLeft / ServiceCountLeftGPU with weights: out[i] = the weight
 * of the keys < cut (ORB_CMP_LT, found cells skipped, out[i] untouched) or
 * <= cut (ORB_CMP_LE, every cell).  Returns nCells * 8.                  */
int orb_count_left_weight(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, int32_t cmp,
                          uint64_t *out);
/* The trace of a weighted build: [cap][32] left weight per pass (UINT64_MAX
 * where a cell was not counted; for an unweighted build the counts).
 * orb_get_pass_counts refuses (ORB_ERR_STATE) after a weighted build.    */
int orb_get_pass_weights(orb_ctx *ctx, uint64_t *sums, int64_t cap);

/* Batched form: capture the whole build into a hipGraph on first use and
 * replay it afterwards (default on; a build with timing enabled runs
 * eagerly).  The capture is redone when the level mode, comparison,
 * partition or trace setting changes.                                    */
int orb_enable_graph(orb_ctx *ctx, bool on);

/* Autotuner for the batched form (no counterpart in the reference).  It
 * picks launch shapes that leave every result unchanged (each variant is
 * exact): which cells go to the LDS subtree kernel (cells of at most
 * subtree_max_keys keys and 64 leaves; 0 = none; the kernel holds at most
 * 8832 keys, 7680 with particle ids, 5120 with weights), and per level the grid
 * cap of the chunk kernels (chunks per block = chunks / blocks) and the
 * subtree kernel's grid (cells per block).  orb_autotune times trial
 * builds of the current particles (device events per level, the best of
 * `repeats` runs each), keeps the fastest choice per level, restores the
 * particles, and applies the result; the graph is recaptured on the next
 * build.  Levels are independent for the per-level shapes because every
 * variant leaves the same particles for the next level.                 */
typedef struct orb_tuning {
    int32_t subtree_max_keys;     /* < 0: default (the kernel capacity) */
    int32_t nlevels;
    int32_t chunk_blocks[64];     /* [level-1], 0 = default                 */
    int32_t subtree_blocks[64];   /* [level-1], 0 = default (a block/cell)  */
    double default_ms;            /* device ms of a build with the defaults */
    double tuned_ms;              /* device ms of a build with the choices  */
} orb_tuning;
int orb_autotune(orb_ctx *ctx, int32_t levels_mode, int32_t repeats, orb_tuning *out);
int orb_set_tuning(orb_ctx *ctx, const orb_tuning *t);  /* NULL: defaults */
int orb_get_tuning(orb_ctx *ctx, orb_tuning *t);

/* Persisted tuning, keyed by level shape (no counterpart in the reference).
 * A row holds the launch shapes chosen for one level, keyed by the level's
 * depth and its cell-size histogram: floor(log2) of the fewest and of the
 * most keys a cell of the level holds under the count rule (a cell of m
 * keys splits at m * ceil(nLeaf / 2) / nLeaf, orbit.cpp:204-205; the leaf
 * counts split as cell.h:79-80).  orb_get_tune_rows lists the context's current choices
 * (after orb_autotune: with each level's tuned device ms); the rows go to a
 * text file and back with orb_tune_table_write / _read; orb_set_tune_rows
 * applies rows to a context of any n and d: each level takes the row with
 * its cell-size key (the same depth first), or the defaults.
 * orb_level_shape gives a level's key: the number of cells of level `level`
 * (1 = root) in a full build of n keys into d leaves, and the fewest and
 * most keys expected in one of them.  Host-only: no GPU needed.          */
typedef struct orb_tune_row {
    int32_t depth;            /* level, 1 = root                          */
    int32_t log2_min_keys;    /* floor(log2(fewest keys)), -1 for none    */
    int32_t log2_max_keys;    /* floor(log2(most keys)), -1 for none      */
    int32_t chunk_blocks;     /* grid cap of the chunk kernels, 0 = default */
    int32_t subtree_blocks;   /* grid cap of the subtree kernel, 0 = default */
    int32_t subtree_max_keys; /* the build's subtree threshold (< 0: default) */
    double ms;                /* the level's tuned device ms, 0 if unknown */
} orb_tune_row;
int orb_level_shape(int64_t n, int32_t d, int32_t level, int64_t *min_keys, int64_t *max_keys);
int orb_tune_table_write(const char *path, const orb_tune_row *rows, int32_t nrows);
int orb_tune_table_read(const char *path, orb_tune_row *rows, int32_t cap);  /* rows read */
int orb_get_tune_rows(orb_ctx *ctx, orb_tune_row *rows, int32_t cap);        /* rows written */
int orb_set_tune_rows(orb_ctx *ctx, const orb_tune_row *rows, int32_t nrows); /* levels matched */

/* Kernel timing of the last build (device events), milliseconds.         */
typedef struct orb_timing {
    double count_ms;       /* bisection count/histogram kernels            */
    double partition_ms;   /* partition kernels                            */
    double other_ms;       /* tables, controller, scans                    */
    int64_t count_bytes;   /* algorithmic bytes read by count kernels      */
    int32_t count_launches;
    int32_t pad;
    double swap_ms;        /* the Hoare swap kernel alone (within partition_ms) */
    int32_t swap_launches;
    int32_t pad2;
    int64_t swap_pairs;    /* pairs the Hoare swaps exchanged (K summed over cells) */
    int64_t swap_bytes;    /* their algorithmic bytes: each pair's two rows read and
                              written, 12 B each (x, y, z) plus 4 B per id and weight */
} orb_timing;
int orb_get_timing(orb_ctx *ctx, orb_timing *t);
/* Per-level time of the last build run with timing on: ms[l-1] for level l
 * (device events from the level's first launch to its last).  Returns the
 * number of levels written.                                              */
int orb_get_level_timing(orb_ctx *ctx, double *ms, int32_t max_levels);
/* HBM bandwidth reference on this GPU (no counterpart in the reference; for
 * reading roofline fractions): a 16-byte streaming read and a copy of the
 * context's particle arrays (3 x n floats), one 16 KB tile per block like
 * the count pass, `reps` timed launches each after an untimed one, in GB/s
 * (copy counts bytes read plus bytes written).  Allocates a scratch copy of
 * the arrays for the duration of the call.                                 */
int orb_measure_stream(orb_ctx *ctx, int32_t reps, double *read_gbs, double *copy_gbs);
int orb_enable_timing(orb_ctx *ctx, bool on);

/* ------------------------------------------------------------------ */
/* Tipsy snapshots: the generate=false input of ServiceInit             */
/* (src/services/init.cu:54-59, whose TipsyIO source is not in the       */
/* reference).  Positions of all particles in file order (gas, dark,     */
/* star).  The standard big-endian (XDR) layout and the native           */
/* little-endian one are told apart by the header.  Host code only.      */
/* ------------------------------------------------------------------ */
typedef struct orb_tipsy_header {
    double time;
    int32_t nbodies, ndim, nsph, ndark, nstar;
    int32_t big_endian;           /* 1: standard XDR file                   */
} orb_tipsy_header;
/* Reads the header; with x, y, z non-NULL also the positions (cap >= nbodies). */
int orb_read_tipsy(const char *path, orb_tipsy_header *hdr, float *x, float *y, float *z,
                   int64_t cap);
/* orb_init(generate=false) with the snapshot's positions as the particles. */
int orb_init_tipsy(orb_ctx **ctx, const char *path, int32_t d, orb_meta_params params);

/* ------------------------------------------------------------------ */
/* synthetic inputs (host code).  No counterpart in the reference beyond */
/* its generator (src/services/init.cu:11-25, 47-53), which orb_generate */
/* extends with a seed; the Plummer sets are the clustered workloads.    */
/* ------------------------------------------------------------------ */
/* xorshf96, i-major draws into SoA arrays: seed 0 is the reference's
 * stream (the particles of orb_init(generate=true)); another seed derives
 * the three state words from splitmix64.                                 */
int orb_generate(uint64_t seed, int64_t n, float *x, float *y, float *z);
/* round(n * background) particles uniform in [-0.5, 0.5)^3, the rest from
 * n_spheres Plummer spheres of scale radius a (radius a / sqrt(u^(-2/3) - 1)
 * for u uniform, isotropic direction), centres uniform in the box, equal
 * split; draws that leave [-0.5, 0.5)^3 as float32 are drawn again.
 * Spheres are spread over the array.  Deterministic in the seed, whatever
 * the number of host threads used.                                       */
int orb_generate_plummer(uint64_t seed, int64_t n, int32_t n_spheres, float a,
                         float background, float *x, float *y, float *z);

/* ------------------------------------------------------------------ */
/* errors / device                                                     */
/* ------------------------------------------------------------------ */
const char *orb_error_string(int code);
int orb_last_hip_error(void);
int orb_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* ORB_H */
