// orbit_binding.cpp -- the reference's master loop (src/orbit.cpp:74-275) on
// the C-ABI, as a maintainer of the reference would bind it.
//
// Compiled against the reference's own src/cell.h (the `Cell` its master
// keeps in `cellHeap`) and include/orb.h, by oracle/Makefile, into the
// git-ignored oracle/_ref/ (it needs the reference tree to build).  Every
// mdl->RunService(PST_X, nIn, vin, vout) of the loop becomes
// orb_run_service(ctx, ORB_PST_X, vin, nIn, vout, nOut) on the same Cell[]
// payloads; the termination rule and the split are the reference's own
// (Cell::getCut, Cell::cut, setCutAxis, setCutMargin from cell.h).  The tree,
// range map and particles it ends with are compared with orb_build on a
// second context holding the same particles, plain and batched forms.
//
//   orbit_binding <log2 N> <log2 d> <o>      (o as in `orbit <x> <y> <o>`)
//
// Prints "binding ok" and exits 0 when everything matches.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <vector>

#include "cell.h"
#include "orb.h"

// the drop-in claim: the reference's Cell and the C-ABI's orb_cell are one layout
static_assert(sizeof(Cell) == sizeof(orb_cell), "Cell size");
static_assert(alignof(Cell) == alignof(orb_cell), "Cell alignment");
static_assert(offsetof(Cell, id) == offsetof(orb_cell, id), "id");
static_assert(offsetof(Cell, nLeafCells) == offsetof(orb_cell, nLeafCells), "nLeafCells");
static_assert(offsetof(Cell, prevCutAxis) == offsetof(orb_cell, prevCutAxis), "prevCutAxis");
static_assert(offsetof(Cell, cutAxis) == offsetof(orb_cell, cutAxis), "cutAxis");
static_assert(offsetof(Cell, foundCut) == offsetof(orb_cell, foundCut), "foundCut");
static_assert(offsetof(Cell, cutMarginLeft) == offsetof(orb_cell, cutMarginLeft), "cutMarginLeft");
static_assert(offsetof(Cell, cutMarginRight) == offsetof(orb_cell, cutMarginRight), "cutMarginRight");
static_assert(offsetof(Cell, lower) == offsetof(orb_cell, lower), "lower");
static_assert(offsetof(Cell, upper) == offsetof(orb_cell, upper), "upper");

namespace {

constexpr int kMaxCells = 8096;  // src/constants.h:11

int fail(const char *what, int rc)
{
    std::printf("FAILED %s: %d (%s)\n", what, rc, orb_error_string(rc));
    return 1;
}

#define CALL(expr)                                  \
    do {                                            \
        int rc_ = (expr);                           \
        if (rc_ < 0) return fail(#expr, rc_);       \
    } while (0)

// orbit.cpp:74-275 with the services on the C-ABI (FAST_MEDIAN is false in
// every configuration orbit.cpp:48-66 sets up).
int master(orb_ctx *ctx, int d, orb_meta_params params, std::vector<Cell> &cellHeap)
{
    float lower[3] = {-0.5, -0.5, -0.5};
    float upper[3] = {0.5, 0.5, 0.5};
    Cell root(0, d, lower, upper);
    root.cutAxis = 0;
    root.setCutMargin();
    cellHeap.assign((size_t)root.getTotalNumberOfCells(), Cell{});
    std::memset(cellHeap.data(), 0, sizeof(Cell) * cellHeap.size());
    cellHeap[0] = root;
    if (params.GPU_PARTITION) CALL(orb_run_service(ctx, ORB_PST_COPYPARTICLES, nullptr, 0, nullptr, 0));

    static unsigned int oCounts[kMaxCells];
    static unsigned int oCountsLeft[kMaxCells];
    for (int l = 1; l < root.getNLevels(); ++l) {
        int a = std::pow(2, (l - 1)) - 1;
        int b = std::min(root.getNCellsOnLastLevel(), (int)std::pow(2, l)) - 2;
        int nCells = b - a + 1;
        Cell *cells = &cellHeap[(size_t)a];
        const int nIn = nCells * (int)sizeof(Cell);
        const int nOut = (int)sizeof oCounts;
        if (!params.GPU_PARTITION)
            CALL(orb_run_service(ctx, ORB_PST_MAKEAXIS, cells, nIn, nullptr, 0));
        CALL(orb_run_service(ctx, ORB_PST_COUNT, cells, nIn, oCounts, nOut));
        if (params.GPU_COUNT && !params.GPU_PARTITION)
            CALL(orb_run_service(ctx, ORB_PST_COPYPARTICLES, nullptr, 0, nullptr, 0));
        if (params.GPU_COUNT) CALL(orb_run_service(ctx, ORB_PST_COPYCELLS, cells, nIn, nullptr, 0));

        bool foundAll = false;
        int j = 0;
        while (!foundAll && j < 32) {
            j++;
            foundAll = true;
            const int srv = params.GPU_PARTITION ? ORB_PST_COUNTLEFTGPU
                            : params.GPU_COUNT   ? ORB_PST_COUNTLEFTAXISGPU
                                                 : ORB_PST_COUNTLEFT;
            CALL(orb_run_service(ctx, srv, cells, nIn, oCountsLeft, nOut));
            for (int i = 0; i < nCells; ++i) {
                if (cells[i].foundCut) continue;
                float ratio = ceil(cells[i].nLeafCells / 2.0) / cells[i].nLeafCells;
                int difference = oCountsLeft[i] - oCounts[i] * ratio;
                if (abs(difference) < 3) {
                    cells[i].foundCut = true;
                } else if (difference > 0) {
                    cells[i].cutMarginRight = cells[i].getCut();
                    foundAll = false;
                } else {
                    cells[i].cutMarginLeft = cells[i].getCut();
                    foundAll = false;
                }
            }
        }
        for (int i = 0; i < nCells; ++i) {
            Cell cellLeft;
            Cell cellRight;
            std::tie(cellLeft, cellRight) = cells[i].cut();
            cellRight.setCutAxis();
            cellRight.setCutMargin();
            cellLeft.setCutAxis();
            cellLeft.setCutMargin();
            cellHeap[(size_t)cells[i].getLeftChildId()] = cellLeft;
            cellHeap[(size_t)cells[i].getRightChildId()] = cellRight;
        }
        CALL(orb_run_service(ctx, params.GPU_PARTITION ? ORB_PST_PARTITIONGPU : ORB_PST_PARTITION,
                             cells, nIn, nullptr, 0));
    }
    return 0;
}

bool same_bits(const void *a, const void *b, size_t n) { return std::memcmp(a, b, n) == 0; }

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::printf("usage: %s <log2 N> <log2 d> <o>\n", argv[0]);
        return 2;
    }
    const int pN = std::atoi(argv[1]), pd = std::atoi(argv[2]), o = std::atoi(argv[3]);
    const int64_t N = int64_t(1) << pN;
    const int d = 1 << pd;
    orb_meta_params params{o >= 1, o >= 2, false};

    orb_ctx *ctx = nullptr;
    CALL(orb_init(&ctx, N, d, true, params));  // PST_INIT (orbit.cpp:83-85)
    std::vector<Cell> cellHeap;
    if (master(ctx, d, params, cellHeap)) return 1;
    const int64_t cap = orb_heap_capacity(d);
    std::vector<uint32_t> ranges((size_t)(2 * cap));
    std::vector<float> x((size_t)N), y((size_t)N), z((size_t)N);
    CALL(orb_get_ranges(ctx, ranges.data(), cap));
    CALL(orb_get_particles(ctx, x.data(), y.data(), z.data(), N));
    CALL(orb_finalize(ctx));  // PST_FINALIZE (orbit.cpp:281)

    int bad = 0;
    for (int form : {ORB_FORM_PLAIN, ORB_FORM_BATCHED}) {
        orb_ctx *c2 = nullptr;
        CALL(orb_init(&c2, N, d, true, params));
        orb_build_info info;
        CALL(orb_build(c2, form, ORB_LEVELS_REFERENCE, &info));
        std::vector<orb_cell> heap((size_t)cap);
        std::vector<uint8_t> exists((size_t)cap);
        std::vector<uint32_t> r2((size_t)(2 * cap));
        std::vector<float> x2((size_t)N), y2((size_t)N), z2((size_t)N);
        CALL(orb_get_heap(c2, heap.data(), cap));
        CALL(orb_get_exists(c2, exists.data(), cap));
        CALL(orb_get_ranges(c2, r2.data(), cap));
        CALL(orb_get_particles(c2, x2.data(), y2.data(), z2.data(), N));
        CALL(orb_finalize(c2));
        int cells_checked = 0;
        for (size_t k = 0; k < cellHeap.size(); ++k) {
            if (!exists[k]) continue;
            ++cells_checked;
            // field by field (the bool's padding bytes carry no value)
            const Cell &a = cellHeap[k];
            const orb_cell &b = heap[k];
            const bool eq = a.id == b.id && a.nLeafCells == b.nLeafCells &&
                            a.prevCutAxis == b.prevCutAxis && a.cutAxis == b.cutAxis &&
                            a.foundCut == b.foundCut &&
                            same_bits(&a.cutMarginLeft, &b.cutMarginLeft, 4) &&
                            same_bits(&a.cutMarginRight, &b.cutMarginRight, 4) &&
                            same_bits(a.lower, b.lower, 12) && same_bits(a.upper, b.upper, 12);
            if (!eq) {
                if (bad < 10) std::printf("form %d: cell %zu differs\n", form, k);
                ++bad;
            }
        }
        if (!same_bits(ranges.data(), r2.data(), 4 * ranges.size())) {
            std::printf("form %d: ranges differ\n", form);
            ++bad;
        }
        if (!same_bits(x.data(), x2.data(), 4 * (size_t)N) ||
            !same_bits(y.data(), y2.data(), 4 * (size_t)N) ||
            !same_bits(z.data(), z2.data(), 4 * (size_t)N)) {
            std::printf("form %d: particles differ\n", form);
            ++bad;
        }
        std::printf("form %d: %d cells compared, %d levels, %d passes\n", form, cells_checked,
                    info.nlevels, info.total_passes);
    }
    if (bad) return 1;
    std::printf("binding ok\n");
    return 0;
}
