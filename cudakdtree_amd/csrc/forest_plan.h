// forest_plan.h -- host plan of a build over many independent trees
// stored back to back in one array (a single tree is a forest of one).
#pragma once

#include <stddef.h>
#include <vector>

#include "cukd_common.h"

namespace cukd_amd {

// a tree on one of its top (radix-select) levels
struct LvlTree {
  int n;       // points in the tree
  int base;    // first element of the tree in the forest arrays
  int seg;     // state index of the tree's segment 0 on this level
  int cseg;    // state index of its segment 0 on the next level
  int hseg;    // first histogram / cursor slot on this level
  int nhseg;   // the same on the next level (when 'more')
  int stile0;  // first select tile
  int ptile0;  // first partition tile
  int tsel;    // select tile size
  int tpart;   // partition tile size
  int more;    // the tree has another top level after this one
};

// a tree entering its LDS-finish levels
struct FinTree {
  int n, base, L0;
  int seg;     // state index of its segment 0 on level L0
  int blk0;    // first workgroup
};

// a tree in the bounds and copy-in passes
struct InTree {
  int n, base, id;
  int seg;     // state index of its root segment
  int tile0;   // first tile of IN_TILE elements
};

enum { IN_TILE = 4096 };

struct ForestPlan {
  int ntrees = 0;
  long long ntotal = 0;
  long long max_n = 0;  // largest tree
  cukd_layout L{};
  bool expl = false;
  bool tables = false;  // tables in device memory (forests) or inline (one tree)
  int max_L0 = 0;
  std::vector<std::vector<LvlTree>> lvl;  // per top level: the trees still on it
  std::vector<int> lvl_segs, lvl_stiles, lvl_ptiles;
  std::vector<int> lvl_hcopies;  // digit histogram copies per subtree, per top level
  std::vector<FinTree> fin[5];            // per finish capacity 512 .. 8192
  int fin_blocks[5] = {0, 0, 0, 0, 0};
  std::vector<InTree> in;
  int in_tiles = 0;
  size_t nstates = 0, max_hseg = 0;
  // workspace layout (bytes)
  size_t off_dataA = 0, off_idxA = 0, off_keyA = 0, off_dataB = 0, off_idxB = 0, off_keyB = 0;
  size_t off_st = 0, off_hist = 0, off_stfin = 0, off_rk = 0, off_cursors = 0, off_tb = 0,
         zero_end = 0, off_acc = 0, off_redo = 0;
  size_t off_tab_in = 0, off_tab_fin[5] = {0, 0, 0, 0, 0};
  std::vector<size_t> off_tab_lvl;
  size_t total = 0;
};

// off: ntrees+1 element offsets (host); device_tables: the plan's tables
// are copied into the workspace (forest_tables_image) instead of being
// passed inline (one tree only)
int make_forest_plan(ForestPlan &P, const long long *off, int ntrees, const cukd_layout &L,
                     bool device_tables);
void forest_tables_image(const ForestPlan &P, std::vector<uint8_t> &img, size_t &begin);
int forest_build_run(const ForestPlan &P, void *d_points, void *d_bounds, uint8_t *ws,
                     hipStream_t s);

}  // namespace cukd_amd
