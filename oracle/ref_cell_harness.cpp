// ref_cell_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference's own src/cell.h (std-only C++) into oracle/_ref so the
// oracle's restated Cell rules can be checked against the original header.
// Built by oracle/Makefile only where the reference tree is present; the
// include path points at it, no reference source is copied here.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include "cell.h"

static_assert(sizeof(Cell) == 52, "Cell layout");

extern "C" {

int ref_cell_size(void) { return (int)sizeof(Cell); }

void ref_make(int id, int nLeafCells, const float *lower, const float *upper, Cell *out)
{
    float lo[3] = {lower[0], lower[1], lower[2]};
    float up[3] = {upper[0], upper[1], upper[2]};
    std::memset(out, 0, sizeof(Cell));
    Cell c(id, nLeafCells, lo, up);
    *out = c;
}

float ref_get_cut(const Cell *c) { return c->getCut(); }

int ref_left_child_id(int id)
{
    Cell c;
    c.id = id;
    return c.getLeftChildId();
}

int ref_right_child_id(int id)
{
    Cell c;
    c.id = id;
    return c.getRightChildId();
}

int ref_n_levels(int d)
{
    Cell c;
    c.nLeafCells = d;
    return c.getNLevels();
}

int ref_n_cells_on_last_level(int d)
{
    Cell c;
    c.nLeafCells = d;
    return c.getNCellsOnLastLevel();
}

int ref_total_cells(int d)
{
    Cell c;
    c.nLeafCells = d;
    return c.getTotalNumberOfCells();
}

void ref_set_cut_axis(Cell *c) { c->setCutAxis(); }

// The split sequence of orbit.cpp:235-243: cut(), then setCutAxis/setCutMargin
// on the right child and then on the left child.  Returns -1 (and writes
// nothing) if a child would get cutAxis == -1, where setCutMargin is undefined.
int ref_split(const Cell *parent, Cell *left, Cell *right)
{
    Cell p = *parent;
    Cell l, r;
    std::tie(l, r) = p.cut();
    r.setCutAxis();
    l.setCutAxis();
    if (r.cutAxis < 0 || l.cutAxis < 0) return -1;
    r.setCutMargin();
    l.setCutMargin();
    std::memset(left, 0, sizeof(Cell));
    std::memset(right, 0, sizeof(Cell));
    *left = l;
    *right = r;
    return 0;
}

}  // extern "C"
