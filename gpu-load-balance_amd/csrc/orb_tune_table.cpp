// orb_tune_table.cpp -- the autotuner's persisted table (host only).
//
// A level is keyed by its depth and its cell-size histogram, summarised as
// floor(log2) of the fewest and of the most keys one of its cells holds.
// Under the count rule those follow from n and d alone: a cell of m keys and
// nLeaf leaves splits at m * ceil(nLeaf / 2) / nLeaf (orbit.cpp:204-205, to
// within the termination rule's +-2) and its leaf counts split as
// ceil(nLeaf / 2), nLeaf - ceil(nLeaf / 2) (cell.h:79-80); only cells with at
// least two leaves are split, and a level holds the cells of its depth.
//
// File format: text, a comment line, then one row per line:
//   depth log2_min_keys log2_max_keys chunk_blocks subtree_blocks subtree_max_keys ms
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "orb.h"

namespace {

struct Node {
    int64_t keys;
    int32_t leaves;
};

}  // namespace

extern "C" {

int orb_level_shape(int64_t n, int32_t d, int32_t level, int64_t *min_keys, int64_t *max_keys)
{
    if (n < 0 || d < 1 || level < 1 || level > 62 || !min_keys || !max_keys) return ORB_ERR_ARG;
    // the cells of depth level-1 with >= 2 leaves, expanding depth by depth;
    // every cell of a depth has one of two leaf counts, so the walk keeps
    // one node per distinct (keys, leaves) pair with its multiplicity
    struct Group {
        Node node;
        int64_t count;
    };
    std::vector<Group> cur = {{{n, d}, 1}};
    for (int32_t l = 1; l < level; ++l) {
        std::vector<Group> next;
        for (const Group &g : cur) {
            if (g.node.leaves < 2) continue;  // a leaf: not split
            const int32_t left = (g.node.leaves + 1) / 2;
            const int64_t kl = (int64_t)std::floor((double)g.node.keys * left / g.node.leaves);
            const Node a{kl, left}, b{g.node.keys - kl, g.node.leaves - left};
            for (const Node &c : {a, b}) {
                bool merged = false;
                for (Group &h : next)
                    if (h.node.keys == c.keys && h.node.leaves == c.leaves) {
                        h.count += g.count;
                        merged = true;
                        break;
                    }
                if (!merged) next.push_back({c, g.count});
            }
        }
        cur.swap(next);
    }
    int64_t cells = 0, lo = -1, hi = -1;
    for (const Group &g : cur) {
        if (g.node.leaves < 2) continue;  // not split at this level
        cells += g.count;
        lo = lo < 0 || g.node.keys < lo ? g.node.keys : lo;
        hi = g.node.keys > hi ? g.node.keys : hi;
    }
    *min_keys = cells ? lo : 0;
    *max_keys = cells ? hi : 0;
    return cells > 0x7fffffff ? 0x7fffffff : (int)cells;
}

int orb_tune_table_write(const char *path, const orb_tune_row *rows, int32_t nrows)
{
    if (!path || nrows < 0 || (nrows > 0 && !rows)) return ORB_ERR_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return ORB_ERR_ARG;
    std::fprintf(f, "# orb tune table: depth log2_min_keys log2_max_keys chunk_blocks "
                    "subtree_blocks subtree_max_keys ms\n");
    for (int32_t i = 0; i < nrows; ++i) {
        const orb_tune_row &r = rows[i];
        std::fprintf(f, "%d %d %d %d %d %d %.17g\n", r.depth, r.log2_min_keys, r.log2_max_keys,
                     r.chunk_blocks, r.subtree_blocks, r.subtree_max_keys, r.ms);
    }
    const bool ok = std::fflush(f) == 0 && !std::ferror(f);
    return (std::fclose(f) == 0 && ok) ? nrows : ORB_ERR_ARG;
}

int orb_tune_table_read(const char *path, orb_tune_row *rows, int32_t cap)
{
    if (!path || cap < 0 || (cap > 0 && !rows)) return ORB_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return ORB_ERR_ARG;
    char line[512];
    int32_t n = 0;
    int rc = 0;
    while (std::fgets(line, sizeof line, f)) {
        const char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '#' || *p == '\n' || *p == '\0') continue;
        orb_tune_row r;
        std::memset(&r, 0, sizeof r);
        int used = 0;
        if (std::sscanf(p, "%d %d %d %d %d %d %lf %n", &r.depth, &r.log2_min_keys,
                        &r.log2_max_keys, &r.chunk_blocks, &r.subtree_blocks,
                        &r.subtree_max_keys, &r.ms, &used) != 7 ||
            p[used] != '\0' || r.depth < 1 || r.depth > 64 || r.log2_min_keys < -1 ||
            r.log2_max_keys < r.log2_min_keys || r.log2_max_keys > 62 || r.chunk_blocks < 0 ||
            r.subtree_blocks < 0 || !(r.ms >= 0.0)) {
            rc = ORB_ERR_ARG;  // malformed row
            break;
        }
        if (n == cap) {
            rc = ORB_ERR_ARG;  // more rows than the caller's array holds
            break;
        }
        rows[n++] = r;
    }
    if (std::ferror(f)) rc = ORB_ERR_ARG;
    std::fclose(f);
    return rc < 0 ? rc : n;
}

}  // extern "C"
