// k_ptlean.hip -- The inlined path tracer of opaque frames.
#include "kcommon.h"

namespace hdr {

// The path tracer on opaque frames whose pixels stay below RNG output 311 (pathlean.h). Block row
// blockIdx.y renders frame frame_index + blockIdx.y of a launch of a.frames frames. kOcc: waves per
// SIMD the registers are allocated for (2: 256 VGPRs, no spills beyond the stacks; 3: 168 VGPRs,
// spilling, faster once a launch holds several waves per SIMD).
template <bool kCount, bool kDebug, int kOcc>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
pt_lean_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views) {
    HDR_TILE(args, tile_base, n_views)
    PixelStats st = {};
    uint32_t draws = 0;
    const int fo = int(blockIdx.y);
    if (valid)
        pt_lean_pixel<kCount, kDebug>(a, x, y, st, draws, a.frame_index + fo,
                                      a.frames > 1 ? a.frame_rgb + size_t(fo) * size_t(a.W) * size_t(a.H) : nullptr);
    if (kCount) {
        unsigned long long *c = a.counters;
        wave_add(c + 1, st.bounce);
        wave_add(c + 2, st.shadow);
        wave_add(c + 3, st.cand);
        wave_add(c + 4, st.cand_skipped);
        wave_add(c + 5, st.ts.nodes);
        wave_add(c + 6, st.ts.tris);
        wave_add(c + 7, st.ts.instances);
        wave_add(c + 8, draws);
    }
}

// ---------------------------------------------------------------------------- instantiations
// (launched from hdrgpu.hip, which sees the declarations in kernels.h)
#define HDR_PT_LEAN(C, D)                                                                                        \
    template __global__ void pt_lean_kernel<C, D, 2>(const FrameArgs *__restrict__, const int *__restrict__, int); \
    template __global__ void pt_lean_kernel<C, D, 3>(const FrameArgs *__restrict__, const int *__restrict__, int);
HDR_PT_LEAN(true, true)
HDR_PT_LEAN(true, false)
HDR_PT_LEAN(false, true)
HDR_PT_LEAN(false, false)
#undef HDR_PT_LEAN

}  // namespace hdr
