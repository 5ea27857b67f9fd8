// k_mega.hip -- The integration megakernel (IntegrationPass + fused accumulation).
#include "kcommon.h"

namespace hdr {

// IntegrationPass (integration_pass.cpp:208-243) + fused accumulation for one pixel. fidx: the
// frame index (seed); stage: the frame's sample goes there instead of into the accumulation (a
// launch of several frames, accumulated in frame order afterwards), or null.
template <bool kCount, bool kDebug>
__device__ __forceinline__ void integrate_pixel(const FrameArgs &a, const GRec *gbuf, int x, int y, double *wlane,
                                                PixelStats &st, uint32_t &draws, int fidx, float4 *stage) {
    const int pixel = y * a.W + x;
    const uint32_t seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)fidx * 12345ULL);
    Tracer<kCount, false> tracer{a.s, st};
    HitRec g;
    const bool gHit = load_grec(gbuf + pixel, g);

    // IntegrationPass (integration_pass.cpp:208-243)
    Mt64 rng;
    uint64_t *slot = a.spill ? a.spill + size_t(pixel) * (2 * kMtN) : nullptr;
    rng.init(seed, slot);
    const float px = float(x) + rng.next_float();
    const float py = float(y) + rng.next_float();
    V3 ro, rd;
    camera_ray(a, px, py, &rng, ro, rd);
    const SpecTab tab = make_tab(rng.next_float());
    Trace dbg = {2166136261u, 2166136261u, 0};
    PixelCtx<kCount, kDebug, false> p{a, tab, rng, tracer, st, dbg, wlane, slot ? slot + kMtN : nullptr};
    const Spec L = path_li(p, rd, gHit, g, pixel);
    V3 rgb = spec2rgb(tab, L);
    if (a.has_cam && a.cam.enable_exposure) rgb = rgb * a.exposure_mul;
    const float4 c4 = make_float4(rgb.x, rgb.y, rgb.z, 1.0f);
    if (stage) {
        stage[pixel] = c4;
    } else if (a.firefly) {
        a.frame_rgb[pixel] = c4;
    } else {
        // AccumulationPass (accumulation_pass.cpp:98-106): accum += c; out = accum * invN
        accumulate_sample(a.accum, a.color, a.lum_sum, a.lum_sq, pixel, c4, a.inv_n);
    }
    if (kDebug) {
        hdr_pixel_debug d;
        d.draws = rng.k;
        d.light_hash = dbg.lhash;
        d.select_hash = dbg.shash;
        d.n_light_picks = dbg.picks;
        d.prim_instance = gHit ? g.inst : -1;
        d.prim_triangle = gHit ? g.tri : -1;
        d.prim_t = gHit ? g.depth : 0.0f;
        d.sample_rgb[0] = rgb.x;
        d.sample_rgb[1] = rgb.y;
        d.sample_rgb[2] = rgb.z;
        a.dbg[pixel] = d;
    }
    draws = rng.k;
    flag_rng_overrun(a, !slot && rng.k > 312u);  // outputs from 312 on need the spill slot
}

// IntegrationPass + (fused) AccumulationPass for one 8x8 pixel tile per wave. Block row blockIdx.y
// renders frame frame_index + blockIdx.y of a launch of a.frames frames. Registers for two waves
// per SIMD (256 VGPRs in all; unconstrained it takes more and runs one): its traversals are
// latency-bound, and the second wave hides them once a launch has the waves (a 16-frame 256 x 256
// RIS launch 1.92 -> 1.15 ms per frame, a translucent 512 x 512 path-traced frame 4.02 -> 2.83 ms;
// a single 256 x 256 frame, one wave per SIMD, is unchanged).
template <bool kCount, bool kDebug>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(2)))
integrate_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views) {
    extern __shared__ double wl[];
    HDR_TILE(args, tile_base, n_views)
    PixelStats st = {};
    uint32_t draws = 0;
    const int fo = int(blockIdx.y);
    if (valid)
        integrate_pixel<kCount, kDebug>(a, a.gbuf, x, y, wl + lane, st, draws, a.frame_index + fo,
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
template __global__ void integrate_kernel<true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void integrate_kernel<true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void integrate_kernel<false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void integrate_kernel<false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);

}  // namespace hdr
