// Library-internal entry between the render server (serve.hip) and the
// compositing kernels (composite.hip); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/gsmi.h"

// Forward of a render-server batch of mixed scenes: the two-level multi-view
// forward with view b's SH rows at view_sh[b] (device array of B pointers)
// and its [H, W, 3] image written to view_out[b]; no per-pixel state.  With
// work (device int[8]) and grid_cap > 0 the tiles are taken from per-XCD
// work queues by at most grid_cap workgroups.
int gsmi_internal_render_served(int64_t V, int B, const gsmi_view* views, const gsmi_view* block_views,
                                int n_tiles_total, const float* mean, const float* cov, const float* const* view_sh,
                                int64_t sh_stride_g, int64_t sh_stride_c, const int32_t* sh_index,
                                const float* alpha, const int32_t* tl, const int32_t* br, const int32_t* bstart,
                                const int32_t* bend, const int32_t* bids, int shift, float* const* view_out, int C,
                                float thresh, int exact, int* work, int grid_cap, hipStream_t st);
