// Per-view tables of the system-matrix traversals (sysmat_walk.h), built on the host by
// sysmat.cpp and read by the host builder and the GPU emission (sysmat_gpu.hip).
#pragma once

#include <vector>

#include "sysmat_walk.h"

namespace xr {
namespace sm {

// Entries before the first view's column tables: a traversal may read one boundary
// before its block (an FP32 start index of -1 after rounding; the reference reads there
// too), which then stays inside the allocation.
constexpr int kTablePad = 2;

struct Tables3d {
    int nview = 0, maxline = 0;
    std::vector<SmView> views;   // [V]
    std::vector<float> cax;      // kTablePad + [V][Nu + 1]: column boundaries on the driving axis
    std::vector<float> center;   // kTablePad + [V][Nu]: column centres
    std::vector<float> delta;    // arc: [V][maxline] line starts
};

struct Tables2d {
    int nview = 0, maxline = 0;
    std::vector<SmView2> views;  // [V]
    std::vector<float> dax;      // kTablePad + [V][Nu + 1]: fan detector boundaries on the driving axis
    std::vector<float> delta;    // arc: [V][maxline] line starts
};

void build_tables3d(const xr_volume3d &vol, const xr_geometry3d &g, const float *ang, int nview, Tables3d &t);
void build_tables2d(const xr_volume2d &vol, const xr_geometry2d &g, const float *ang, int nview, Tables2d &t);

}  // namespace sm
}  // namespace xr
