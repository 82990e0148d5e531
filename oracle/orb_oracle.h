/*
 * orb_oracle.h -- CPU restatement of the reference ORB build (o=0 path).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation: only tests, the smoke entry and the CPU-baseline leg of the
 * benchmark may load it.  Nothing in the product links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Compile with -O2 -ffp-contract=off and
 * without -ffast-math: the termination rule and the cut midpoint are float32
 * arithmetic whose results change under FMA contraction.
 */
#ifndef ORB_ORACLE_H
#define ORB_ORACLE_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field list, order and padding as `struct Cell` (src/cell.h:9-17):
 * 52 bytes, 4-byte aligned.                                                 */
typedef struct orc_cell {
    int32_t id;
    int32_t nLeafCells;
    int32_t prevCutAxis;
    int32_t cutAxis;
    bool foundCut;
    float cutMarginLeft;
    float cutMarginRight;
    float lower[3];
    float upper[3];
} orc_cell;

enum {
    ORC_OK = 0,
    ORC_ERR_ARG = -1,
    ORC_ERR_UB = -4,     /* a case where the reference has undefined behaviour */
};

enum { ORC_LEVELS_REFERENCE = 0, ORC_LEVELS_FULL = 1 };
enum { ORC_CMP_LT = 0, ORC_CMP_LE = 1 };
enum { ORC_PART_HOARE = 0, ORC_PART_STABLE_LE = 1 };

#define ORC_MAX_PASSES 32

/* ---- generator (src/services/init.cu:11-25, 47-53) ---- */
typedef struct orc_rng { uint64_t x, y, z; } orc_rng;
void orc_rng_reference_seed(orc_rng *s);
void orc_rng_step(orc_rng *s);                 /* state update only          */
float orc_rng_value(uint64_t z);               /* z -> float in [-0.5, 0.5]  */
float orc_xorshf96(orc_rng *s);                /* step, then value of new z  */
/* i-major draws (x_i, y_i, z_i) stored SoA, fresh reference seeds.         */
void orc_generate(int64_t n, float *x, float *y, float *z);

/* ---- Cell rules (src/cell.h) ---- */
orc_cell orc_cell_make(int32_t id, int32_t nLeafCells, const float lower[3], const float upper[3]);
int32_t orc_left_child_id(int32_t id);
int32_t orc_right_child_id(int32_t id);
int32_t orc_n_levels(int32_t nLeafCells);
int32_t orc_n_cells_on_last_level(int32_t nLeafCells);
float orc_get_cut(const orc_cell *c);
void orc_set_cut_axis(orc_cell *c);
void orc_set_cut_margin(orc_cell *c);
/* cut() + setCutAxis() + setCutMargin() on both children, as orbit.cpp:235-250
 * applies them.  Returns ORC_ERR_UB when a child gets cutAxis == -1.          */
int orc_split(const orc_cell *parent, orc_cell *left, orc_cell *right);
orc_cell orc_root(int32_t d);

/* ---- termination rule (src/orbit.cpp:204-208) ---- */
int32_t orc_difference(uint32_t count_left, uint32_t count, int32_t nLeafCells);
/* The same rule on weight sums (weighted build): the weight left of the cut
 * against the cell's weight times the left share, in the same float
 * roundings; the truncation keeps 64 bits.  With every weight 1 the sums are
 * the counts and the result is orc_difference's.                           */
int64_t orc_difference_w(uint64_t weight_left, uint64_t weight, int32_t nLeafCells);

/* ---- partitions ---- */
/* Serial Hoare loop of src/services/partition.cpp:30-52 on SoA columns,
 * restricted to [begin,end).  *split receives i (absolute index).
 * Returns ORC_ERR_UB for the inputs where the reference reads or swaps
 * outside the cell (empty cell, or no key >= cut).                          */
int orc_hoare(float *x, float *y, float *z, int64_t begin, int64_t end, int axis,
              float cut, int64_t *split);
/* The same permutation from two scans (closed form).  perm[p] receives the
 * source index (cell-local) of the element that ends at position p.        */
int orc_hoare_closed_form(const float *key, int64_t n, float cut, int64_t *K,
                          int64_t *split, int64_t *perm);
/* Stable predicate partition on `<=` (deterministic restatement of the set
 * semantics of src/services/partitionGPU.cu:58-231).                       */
int orc_stable_le(float *x, float *y, float *z, int64_t begin, int64_t end, int axis,
                  float cut, int64_t *split);

/* ---- the full build (src/orbit.cpp:74-275, o=0 service sequence) ---- */
typedef struct orc_opts {
    int32_t levels_mode;  /* ORC_LEVELS_*                                   */
    int32_t cmp;          /* ORC_CMP_*                                      */
    int32_t partition;    /* ORC_PART_*                                     */
    int32_t max_passes;   /* 32 in the reference                            */
    int32_t threads;      /* <= 1: serial; else OpenMP over cells and over
                             the keys of a large cell.  Counts are integer
                             sums and cells are independent, so the result
                             is the same for every thread count.          */
} orc_opts;

typedef struct orc_trace {
    orc_cell *heap;            /* [heap_cap] cells by heap id                 */
    uint32_t *ranges;          /* [heap_cap][2] cellToRangeMap                */
    uint8_t *exists;           /* [heap_cap] 1 if the cell was created        */
    uint32_t *pass_counts;     /* optional [heap_cap][32] left count per pass */
    int32_t *level_passes;     /* [64] passes per level                       */
    int64_t heap_cap;
    int32_t nlevels;           /* levels processed                            */
    int32_t total_passes;
    int32_t error_cell;        /* heap id that raised ORC_ERR_UB, else -1     */
    int32_t pad;
    uint32_t *ids;             /* optional [n]: set to 0..n-1, then moved with
                                  the particles by every partition            */
    uint32_t *weights;         /* optional [n]: the workload weight w(p_i) of
                                  each particle (documentation/notes.tex:259-271),
                                  moved with the particles; NULL: every weight 1 */
    uint64_t *pass_weights;    /* optional [heap_cap][32]: the left weight sum of
                                  each pass (the left count when unweighted)  */
    double count_ms;           /* out: wall time of the counting passes and the
                                  controller (the reference's CountCopy column,
                                  orbit.cpp:164-188)                          */
    double partition_ms;       /* out: wall time of the partitions (its
                                  Partition column, orbit.cpp:252-274)        */
} orc_trace;

/* Capacity of the heap for the given d and level mode (ids are 0-based).   */
int64_t orc_heap_capacity(int32_t d, int32_t levels_mode);
int orc_build(int64_t n, int32_t d, const orc_opts *opts, float *x, float *y, float *z,
              orc_trace *trace);

/* Cells processed on level l (1-based) in the schedule; writes their ids in
 * heap order into ids (capacity cap) and returns the count.                */
int64_t orc_level_cells(int32_t d, int32_t levels_mode, int32_t l, const uint8_t *exists,
                        const orc_cell *heap, int32_t *ids, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
