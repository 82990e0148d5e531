/*
 * orb_oracle.c -- CPU restatement of the reference ORB build (o=0 path).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation and
 * the timed CPU baseline.  Only tests, the smoke entry and the benchmark's CPU
 * baseline leg load it; the product never does.
 *
 * Each function cites the reference file:line it restates.  Build with
 * `-O2 -ffp-contract=off` (no -ffast-math): see oracle/Makefile.
 *
 * Parity pins (what checks each part independently of this file):
 *   generator       against the reference's own xorshf96 (init.cu:11-25,
 *                   compiled where it lies into oracle/_ref) and the fixture
 *                   it wrote (first 2^20 particles); uint64->float32 rounding
 *                   against exact arithmetic; uniform moments.
 *   weights         unit weights give the unweighted build; a particle of
 *                   weight k builds the same tree as k copies of it
 *                   (tests/test_oracle_weights.py).
 *   Cell rules      against oracle/_ref (the reference's own src/cell.h compiled
 *                   by oracle/Makefile) and the committed fixture it wrote.
 *   termination     brute force over small counts against exact rational
 *                   arithmetic with the two float roundings applied by numpy.
 *   Hoare           the serial loop against the closed form and against the
 *                   thesis worked example (documentation/particles10-12.csv).
 *   build           invariants (ranges tile the parent, child sizes, found
 *                   cells satisfy |diff| < 3, cut inside the margins) and a
 *                   brute-force bisection on tiny inputs.
 */
#include "orb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------------ */
/* generator: src/services/init.cu:11-25 (xorshf96 on 64-bit unsigned long)  */
/* ------------------------------------------------------------------------ */

void orc_rng_reference_seed(orc_rng *s)
{
    /* init.cu:11 */
    s->x = 123456789ull;
    s->y = 362436069ull;
    s->z = 521288629ull;
}

void orc_rng_step(orc_rng *s)
{
    /* init.cu:15-22 */
    uint64_t t;
    s->x ^= s->x << 16;
    s->x ^= s->x >> 5;
    s->x ^= s->x << 1;
    t = s->x;
    s->x = s->y;
    s->y = s->z;
    s->z = t ^ s->x ^ s->y;
}

float orc_rng_value(uint64_t z)
{
    /* init.cu:24: (float)z / ULONG_MAX - 0.5, returned as float.
     * ULONG_MAX converts to the float 2^64; the quotient is float; the
     * subtraction of the double 0.5 happens in double; the return narrows. */
    float f = (float)z;
    float q = f / (float)UINT64_MAX;
    double v = (double)q - 0.5;
    return (float)v;
}

float orc_xorshf96(orc_rng *s)
{
    orc_rng_step(s);
    return orc_rng_value(s->z);
}

void orc_generate(int64_t n, float *x, float *y, float *z)
{
    /* init.cu:47-53: for i: for d in 0..2: particles(i,d) = xorshf96();
     * blitz storage ordering (0,1) makes particles(N,3) axis-major SoA.     */
    orc_rng s;
    orc_rng_reference_seed(&s);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = orc_xorshf96(&s);
        y[i] = orc_xorshf96(&s);
        z[i] = orc_xorshf96(&s);
    }
}

/* ------------------------------------------------------------------------ */
/* Cell rules: src/cell.h                                                   */
/* ------------------------------------------------------------------------ */

orc_cell orc_cell_make(int32_t id, int32_t nLeafCells, const float lower[3], const float upper[3])
{
    /* cell.h:19-45 (constructor) */
    orc_cell c;
    memset(&c, 0, sizeof c);
    c.id = id;
    c.nLeafCells = nLeafCells;
    c.foundCut = false;
    c.prevCutAxis = -1;
    c.cutAxis = -1;
    c.cutMarginLeft = 0.0f;
    c.cutMarginRight = 0.0f;
    for (int d = 0; d < 3; ++d) {
        c.lower[d] = lower[d];
        c.upper[d] = upper[d];
    }
    return c;
}

int32_t orc_left_child_id(int32_t id) { return (id + 1) * 2 - 1; }  /* cell.h:49-51 */
int32_t orc_right_child_id(int32_t id) { return (id + 1) * 2; }     /* cell.h:53-55 */

int32_t orc_n_levels(int32_t nLeafCells)
{
    /* cell.h:65-67: ceil(log2(nLeafCells)), exact in integers for d <= 2^30 */
    int32_t D = 0;
    while (((int64_t)1 << D) < (int64_t)nLeafCells) ++D;
    return D;
}

int32_t orc_n_cells_on_last_level(int32_t nLeafCells)
{
    /* cell.h:69-72: 2*nLeafCells - 2^depth */
    return 2 * nLeafCells - (int32_t)((int64_t)1 << orc_n_levels(nLeafCells));
}

float orc_get_cut(const orc_cell *c)
{
    /* cell.h:74-76: float sum, double halving, float return */
    float s = c->cutMarginRight + c->cutMarginLeft;
    return (float)((double)s / 2.0);
}

void orc_set_cut_axis(orc_cell *c)
{
    /* cell.h:102-121: longest side, strict '>', first max wins, -1 if none */
    int maxD = -1;
    float maxSize = 0.0f;
    for (int d = 0; d < 3; d++) {
        float size = c->upper[d] - c->lower[d];
        if (size > maxSize) {
            maxSize = size;
            maxD = d;
        }
    }
    c->cutAxis = maxD;
}

void orc_set_cut_margin(orc_cell *c)
{
    /* cell.h:123-126 (callers check cutAxis >= 0 first) */
    c->cutMarginLeft = c->lower[c->cutAxis];
    c->cutMarginRight = c->upper[c->cutAxis];
}

int orc_split(const orc_cell *parent, orc_cell *left, orc_cell *right)
{
    /* cell.h:78-100 (cut) followed by orbit.cpp:240-243 */
    if (parent->cutAxis < 0 || parent->cutAxis > 2) return ORC_ERR_UB;
    int32_t nLeft = (int32_t)ceil(parent->nLeafCells / 2.0);
    int32_t nRight = parent->nLeafCells - nLeft;
    float cut = orc_get_cut(parent);
    *left = orc_cell_make(orc_left_child_id(parent->id), nLeft, parent->lower, parent->upper);
    left->upper[parent->cutAxis] = cut;
    left->prevCutAxis = parent->cutAxis;
    *right = orc_cell_make(orc_right_child_id(parent->id), nRight, parent->lower, parent->upper);
    right->lower[parent->cutAxis] = cut;
    right->prevCutAxis = parent->cutAxis;
    orc_set_cut_axis(right);
    if (right->cutAxis < 0) return ORC_ERR_UB;
    orc_set_cut_margin(right);
    orc_set_cut_axis(left);
    if (left->cutAxis < 0) return ORC_ERR_UB;
    orc_set_cut_margin(left);
    return ORC_OK;
}

orc_cell orc_root(int32_t d)
{
    /* orbit.cpp:45-46, 74-76 */
    const float lower[3] = {-0.5f, -0.5f, -0.5f};
    const float upper[3] = {0.5f, 0.5f, 0.5f};
    orc_cell r = orc_cell_make(0, d, lower, upper);
    r.cutAxis = 0;
    orc_set_cut_margin(&r);
    return r;
}

/* ------------------------------------------------------------------------ */
/* termination rule: src/orbit.cpp:204-208                                   */
/* ------------------------------------------------------------------------ */

int32_t orc_difference(uint32_t count_left, uint32_t count, int32_t nLeafCells)
{
    /* float ratio = ceil(n/2.0)/n;  int difference = L - C*ratio;
     * the quotient is double narrowed to float; C converts to float, the
     * product and the difference are float; the int conversion truncates.   */
    float ratio = (float)(ceil(nLeafCells / 2.0) / (double)nLeafCells);
    float prod = (float)count * ratio;
    float diff = (float)count_left - prod;
    return (int32_t)diff;
}

int64_t orc_difference_w(uint64_t weight_left, uint64_t weight, int32_t nLeafCells)
{
    /* orbit.cpp:204-205 with the counts replaced by weight sums
     * (notes.tex:326-336 states the weighted split target; the reference's
     * controller reaches it by the same bisection).  uint64 -> float rounds
     * to nearest even, as uint32 -> float does.                             */
    float ratio = (float)(ceil(nLeafCells / 2.0) / (double)nLeafCells);
    float prod = (float)weight * ratio;
    float diff = (float)weight_left - prod;
    return (int64_t)diff;
}

/* ------------------------------------------------------------------------ */
/* partitions                                                               */
/* ------------------------------------------------------------------------ */

/* payload that moves with the rows: particle ids and weights (optional) */
typedef struct rows_payload { uint32_t *ids, *w; } rows_payload;

static void swap_rows(float *x, float *y, float *z, rows_payload pl, int64_t a, int64_t b)
{
    /* partition.cpp:10-16; ids and weights (optional) move with the row */
    float t;
    t = x[a]; x[a] = x[b]; x[b] = t;
    t = y[a]; y[a] = y[b]; y[b] = t;
    t = z[a]; z[a] = z[b]; z[b] = t;
    if (pl.ids) { uint32_t u = pl.ids[a]; pl.ids[a] = pl.ids[b]; pl.ids[b] = u; }
    if (pl.w) { uint32_t u = pl.w[a]; pl.w[a] = pl.w[b]; pl.w[b] = u; }
}

static int hoare_impl(float *x, float *y, float *z, rows_payload pl, int64_t begin, int64_t end,
                      int axis, float cut, int64_t *split);
static int stable_impl(float *x, float *y, float *z, rows_payload pl, int64_t begin, int64_t end,
                       int axis, float cut, int64_t *split);

int orc_hoare(float *x, float *y, float *z, int64_t begin, int64_t end, int axis,
              float cut, int64_t *split)
{
    rows_payload none = {NULL, NULL};
    return hoare_impl(x, y, z, none, begin, end, axis, cut, split);
}

int orc_stable_le(float *x, float *y, float *z, int64_t begin, int64_t end, int axis,
                  float cut, int64_t *split)
{
    rows_payload none = {NULL, NULL};
    return stable_impl(x, y, z, none, begin, end, axis, cut, split);
}

static int hoare_impl(float *x, float *y, float *z, rows_payload pl, int64_t begin, int64_t end,
                      int axis, float cut, int64_t *split)
{
    /* partition.cpp:28-60.  The reference's bounds tests come after the key
     * read (`p(i) < cut && i <= end`, `p(j) > cut && j >= begin`); they are
     * evaluated first here, which gives the same indices whenever a key >= cut
     * exists in the cell (checked below) and avoids the out-of-cell read.    */
    float *k = axis == 0 ? x : axis == 1 ? y : z;
    if (axis < 0 || axis > 2 || end <= begin) return ORC_ERR_UB;
    int any = 0;
    for (int64_t p = begin; p < end; ++p)
        if (k[p] >= cut) { any = 1; break; }
    if (!any) return ORC_ERR_UB;  /* the reference swaps outside the cell */

    int64_t i = begin - 1, j = end;
    for (;;) {
        do { i++; } while (i < end && k[i] < cut);
        do { j--; } while (j >= begin && k[j] > cut);
        if (i >= j) break;
        swap_rows(x, y, z, pl, i, j);
    }
    swap_rows(x, y, z, pl, i, end - 1);
    *split = i;
    return ORC_OK;
}

int orc_hoare_closed_form(const float *key, int64_t n, float cut, int64_t *K_out,
                          int64_t *split, int64_t *perm)
{
    /* Hoare as scans.  L-stoppers (key >= cut) left to right L_0 < L_1 < ...;
     * R-stoppers (key <= cut) right to left R_0 > R_1 > ...;  K = #{t : L_t < R_t}.
     * The loop swaps (L_t, R_t) for t < K and stops with
     * i = min(L_K, R_{K-1}) (i = L_0 when K == 0); then swaps (i, n-1).      */
    if (n <= 0) return ORC_ERR_UB;
    int64_t *Lp = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *Rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t TL = 0, TR = 0;
    for (int64_t p = 0; p < n; ++p)
        if (key[p] >= cut) Lp[TL++] = p;
    for (int64_t p = n - 1; p >= 0; --p)
        if (key[p] <= cut) Rp[TR++] = p;
    if (TL == 0) { free(Lp); free(Rp); return ORC_ERR_UB; }
    int64_t K = 0;
    while (K < TL && K < TR && Lp[K] < Rp[K]) ++K;
    int64_t i;
    if (K == 0) {
        i = Lp[0];
    } else {
        i = Rp[K - 1];
        if (K < TL && Lp[K] < i) i = Lp[K];
    }
    for (int64_t p = 0; p < n; ++p) perm[p] = p;
    for (int64_t t = 0; t < K; ++t) {
        int64_t a = Lp[t], b = Rp[t];
        int64_t s = perm[a]; perm[a] = perm[b]; perm[b] = s;
    }
    {
        int64_t s = perm[i]; perm[i] = perm[n - 1]; perm[n - 1] = s;
    }
    *K_out = K;
    *split = i;
    free(Lp);
    free(Rp);
    return ORC_OK;
}

static int stable_impl(float *x, float *y, float *z, rows_payload pl, int64_t begin, int64_t end,
                       int axis, float cut, int64_t *split)
{
    /* Keys <= cut first, then keys > cut, each in input order.  The reference
     * o=2 kernel (partitionGPU.cu:58-231) produces these two sets in a
     * nondeterministic order; this stable order is the deterministic choice. */
    if (axis < 0 || axis > 2 || end < begin) return ORC_ERR_ARG;
    int64_t n = end - begin;
    float *k = axis == 0 ? x : axis == 1 ? y : z;
    float *tmp = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    uint32_t *ids = pl.ids, *wts = pl.w;
    uint32_t *tid = ids ? (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1)) : NULL;
    uint32_t *tw = wts ? (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1)) : NULL;
    int64_t nl = 0;
    for (int64_t p = begin; p < end; ++p)
        if (k[p] <= cut) nl++;
    int64_t wl = 0, wr = nl;
    for (int64_t p = begin; p < end; ++p) {
        int64_t w = k[p] <= cut ? wl++ : wr++;
        tmp[w] = x[p];
        tmp[n + w] = y[p];
        tmp[2 * n + w] = z[p];
        if (tid) tid[w] = ids[p];
        if (tw) tw[w] = wts[p];
    }
    memcpy(x + begin, tmp, sizeof(float) * (size_t)n);
    memcpy(y + begin, tmp + n, sizeof(float) * (size_t)n);
    memcpy(z + begin, tmp + 2 * n, sizeof(float) * (size_t)n);
    if (tid) memcpy(ids + begin, tid, sizeof(uint32_t) * (size_t)n);
    if (tw) memcpy(wts + begin, tw, sizeof(uint32_t) * (size_t)n);
    free(tmp);
    free(tid);
    free(tw);
    *split = begin + nl;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* level schedule: src/orbit.cpp:102-111                                     */
/* ------------------------------------------------------------------------ */

int64_t orc_heap_capacity(int32_t d, int32_t levels_mode)
{
    (void)levels_mode;
    if (d < 1) return 0;
    int32_t D = orc_n_levels(d);
    int64_t cap = ((int64_t)1 << (D + 1)) - 1;
    int64_t ref = 2 * (int64_t)d - 1;  /* cell.h:61-63 */
    return cap > ref ? cap : ref;
}

int64_t orc_level_cells(int32_t d, int32_t levels_mode, int32_t l, const uint8_t *exists,
                        const orc_cell *heap, int32_t *ids, int64_t cap)
{
    int32_t D = orc_n_levels(d);
    if (l < 1) return 0;
    int64_t cnt = 0;
    if (levels_mode == ORC_LEVELS_REFERENCE) {
        /* orbit.cpp:102-110: l < getNLevels(); a = 2^(l-1)-1;
         * b = min(getNCellsOnLastLevel(), 2^l) - 2.                          */
        if (l >= D) return 0;
        int64_t a = ((int64_t)1 << (l - 1)) - 1;
        int64_t last = orc_n_cells_on_last_level(d);
        int64_t p2 = (int64_t)1 << l;
        int64_t b = (last < p2 ? last : p2) - 2;
        for (int64_t id = a; id <= b; ++id) {
            if (cnt < cap) ids[cnt] = (int32_t)id;
            cnt++;
        }
        return cnt;
    }
    /* full tree: every existing cell of depth l-1 with more than one leaf */
    if (l > D) return 0;
    int64_t a = ((int64_t)1 << (l - 1)) - 1;
    int64_t b = ((int64_t)1 << l) - 2;
    for (int64_t id = a; id <= b; ++id) {
        if (!exists[id] || heap[id].nLeafCells < 2) continue;
        if (cnt < cap) ids[cnt] = (int32_t)id;
        cnt++;
    }
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* the build: src/orbit.cpp:74-275 with the o=0 services                     */
/* ------------------------------------------------------------------------ */

int orc_build(int64_t n, int32_t d, const orc_opts *opts, float *x, float *y, float *z,
              orc_trace *tr)
{
    if (n < 0 || n > 0xffffffffll || d < 1 || !opts || !tr) return ORC_ERR_ARG;
    int64_t cap = orc_heap_capacity(d, opts->levels_mode);
    if (tr->heap_cap < cap || !tr->heap || !tr->ranges || !tr->exists || !tr->level_passes)
        return ORC_ERR_ARG;
    int32_t max_passes = opts->max_passes > 0 ? opts->max_passes : ORC_MAX_PASSES;
    if (max_passes > ORC_MAX_PASSES) return ORC_ERR_ARG;

    memset(tr->heap, 0, sizeof(orc_cell) * (size_t)cap);
    memset(tr->ranges, 0, sizeof(uint32_t) * 2 * (size_t)cap);
    memset(tr->exists, 0, (size_t)cap);
    if (tr->pass_counts)
        memset(tr->pass_counts, 0xff, sizeof(uint32_t) * ORC_MAX_PASSES * (size_t)cap);
    for (int i = 0; i < 64; ++i) tr->level_passes[i] = 0;
    tr->nlevels = 0;
    tr->total_passes = 0;
    tr->error_cell = -1;
    tr->count_ms = 0.0;
    tr->partition_ms = 0.0;

    if (tr->ids)
        for (int64_t p = 0; p < n; ++p) tr->ids[p] = (uint32_t)p;
    tr->heap[0] = orc_root(d);
    tr->exists[0] = 1;
    tr->ranges[0] = 0;                   /* init.cu:63-66 */
    tr->ranges[1] = (uint32_t)n;

    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    /* weight sums; with no weights given every weight is 1 and these are the
     * counts of count.cpp:14-17 and countLeft.cpp:33-36                     */
    uint64_t *counts = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cap);
    uint64_t *left = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cap);
    const uint32_t *wt = tr->weights;
    rows_payload pl = {tr->ids, tr->weights};
    if (tr->pass_weights)
        memset(tr->pass_weights, 0xff, sizeof(uint64_t) * ORC_MAX_PASSES * (size_t)cap);
    const int nt = opts->threads > 1 ? opts->threads : 1;
    int rc = ORC_OK;
    int32_t D = orc_n_levels(d);

    for (int32_t l = 1; l <= D; ++l) {
        int64_t nCells = orc_level_cells(d, opts->levels_mode, l, tr->exists, tr->heap, ids, cap);
        if (opts->levels_mode == ORC_LEVELS_REFERENCE && l >= D) break;
        tr->nlevels = l;
        if (nCells <= 0) continue;
        for (int64_t i = 0; i < nCells; ++i) {
            /* the reference schedule can name ids that were never written
             * (non-power-of-two d); it then reads uninitialised cells        */
            if (!tr->exists[ids[i]]) { tr->error_cell = ids[i]; rc = ORC_ERR_UB; goto done; }
        }

        /* ServiceCount (count.cpp:14-17); weighted: the cell's weight */
        const double t_count = now_ms();
        for (int64_t i = 0; i < nCells; ++i) {
            int32_t id = ids[i];
            uint32_t b = tr->ranges[2 * id], e = tr->ranges[2 * id + 1];
            if (!wt) {
                counts[i] = e - b;
            } else {
                uint64_t s = 0;
                for (uint32_t p = b; p < e; ++p) s += wt[p];
                counts[i] = s;
            }
        }

        /* bisection controller (orbit.cpp:146-232).  The counts of a pass
         * come first (threads over cells, or over the keys of a cell when a
         * level has fewer cells than threads), then the serial update.     */
        int foundAll = 0;
        int32_t j = 0;
        while (!foundAll && j < max_passes) {
            j++;
            foundAll = 1;
            const int by_cells = nt > 1 && nCells >= nt;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) if (by_cells)
            for (int64_t i = 0; i < nCells; ++i) {
                const orc_cell *c = &tr->heap[ids[i]];
                if (c->foundCut) continue;  /* countLeft.cpp:19-21 */
                uint32_t b = tr->ranges[2 * c->id], e = tr->ranges[2 * c->id + 1];
                const float *col = c->cutAxis == 0 ? x : c->cutAxis == 1 ? y : z;
                float cut = orc_get_cut(c);
                uint64_t nLeft = 0;
                const int by_keys = nt > 1 && !by_cells && (e - b) > 65536u;
                if (!wt && opts->cmp == ORC_CMP_LT) {
#pragma omp parallel for reduction(+ : nLeft) num_threads(nt) if (by_keys)
                    for (uint32_t p = b; p < e; ++p) nLeft += col[p] < cut;   /* countLeft.cpp:33-36 */
                } else if (!wt) {
#pragma omp parallel for reduction(+ : nLeft) num_threads(nt) if (by_keys)
                    for (uint32_t p = b; p < e; ++p) nLeft += col[p] <= cut;  /* countLeftGPUAxis.cu:159-162 */
                } else if (opts->cmp == ORC_CMP_LT) {
#pragma omp parallel for reduction(+ : nLeft) num_threads(nt) if (by_keys)
                    for (uint32_t p = b; p < e; ++p) nLeft += col[p] < cut ? wt[p] : 0u;
                } else {
#pragma omp parallel for reduction(+ : nLeft) num_threads(nt) if (by_keys)
                    for (uint32_t p = b; p < e; ++p) nLeft += col[p] <= cut ? wt[p] : 0u;
                }
                left[i] = nLeft;
            }
            for (int64_t i = 0; i < nCells; ++i) {
                orc_cell *c = &tr->heap[ids[i]];
                if (c->foundCut) continue;
                const uint64_t nLeft = left[i];
                const int64_t slot = (int64_t)c->id * ORC_MAX_PASSES + (j - 1);
                if (tr->pass_counts && !wt) tr->pass_counts[slot] = (uint32_t)nLeft;
                if (tr->pass_weights) tr->pass_weights[slot] = nLeft;
                int64_t difference = wt ? orc_difference_w(nLeft, counts[i], c->nLeafCells)
                                        : orc_difference((uint32_t)nLeft, (uint32_t)counts[i],
                                                         c->nLeafCells);
                if (difference < 3 && difference > -3) {
                    c->foundCut = true;
                } else if (difference > 0) {
                    c->cutMarginRight = orc_get_cut(c);
                    foundAll = 0;
                } else {
                    c->cutMarginLeft = orc_get_cut(c);
                    foundAll = 0;
                }
            }
        }
        tr->level_passes[l - 1] = j;
        tr->total_passes += j;
        tr->count_ms += now_ms() - t_count;

        /* split (orbit.cpp:235-250) */
        for (int64_t i = 0; i < nCells; ++i) {
            orc_cell *c = &tr->heap[ids[i]];
            orc_cell left, right;
            int r = orc_split(c, &left, &right);
            if (r != ORC_OK) { tr->error_cell = c->id; rc = r; goto done; }
            tr->heap[left.id] = left;
            tr->heap[right.id] = right;
            tr->exists[left.id] = 1;
            tr->exists[right.id] = 1;
        }

        /* partition (partition.cpp:18-60); cells are disjoint ranges     */
        int bad = -1;
        const double t_part = now_ms();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) if (nt > 1 && nCells > 1)
        for (int64_t i = 0; i < nCells; ++i) {
            orc_cell *c = &tr->heap[ids[i]];
            int64_t b = tr->ranges[2 * c->id], e = tr->ranges[2 * c->id + 1];
            int64_t split = 0;
            int r;
            if (opts->partition == ORC_PART_HOARE)
                r = hoare_impl(x, y, z, pl, b, e, c->cutAxis, orc_get_cut(c), &split);
            else
                r = stable_impl(x, y, z, pl, b, e, c->cutAxis, orc_get_cut(c), &split);
            if (r != ORC_OK) {
#pragma omp critical
                {
                    if (bad < 0 || i < bad) { bad = (int)i; rc = r; }
                }
                continue;
            }
            int32_t lid = orc_left_child_id(c->id), rid = orc_right_child_id(c->id);
            tr->ranges[2 * lid] = (uint32_t)b;
            tr->ranges[2 * lid + 1] = (uint32_t)split;
            tr->ranges[2 * rid] = (uint32_t)split;
            tr->ranges[2 * rid + 1] = (uint32_t)e;
        }
        tr->partition_ms += now_ms() - t_part;
        if (bad >= 0) { tr->error_cell = tr->heap[ids[bad]].id; goto done; }
    }
done:
    free(ids);
    free(counts);
    free(left);
    return rc;
}
