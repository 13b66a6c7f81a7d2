// Host-side layout of a batch of scenes, shared by the batched system
// (batch_system.cpp) and the batched adjoint (batch_adjoint_system.cpp).
#pragma once

#include <vector>

#include "batch_types.h"
#include "rtps_internal.h"

namespace rtps {

struct BatchLayout {
    std::vector<RTPSSettings> settings;
    std::vector<SceneDev> scenes;   // offsets: scene capacities laid end to end
    std::vector<BatchItem> items;   // 256-slot tiles of every scene, in scene order
    std::vector<BatchItem> items128;   // the same in 128-slot tiles
    std::vector<BatchItem> items64;    // and in 64-slot tiles
    std::vector<int> item_first;    // first item of scene s (num_scenes + 1 entries)
    long T = 0, total_cells = 0;
    int shifts[4] = {}, widths[4] = {}, npasses = 0;   // sort plan of the segmented keys
};

// Parses the scenes' settings (n_settings[s] key/value pairs each, laid end
// to end), derives every scene's parameters and grid, and plans the sort and
// the work items.  Returns an RTPS_* status (with rtps_last_error set).
int build_batch_layout(int num_scenes, const int* n_settings, const char* const* keys, const char* const* vals,
                       BatchLayout& out);

}  // namespace rtps
