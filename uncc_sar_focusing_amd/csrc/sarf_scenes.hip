// sarf_scenes.hip -- backprojection of many independent scenes of uneven
// size (pulses, profile length, image size, geometry) in one persistent
// launch.
//
// The reference focuses one aperture per call (cuda_focus_SAR_image,
// cuda_sar_focusing.hpp:526-723), allocating and resetting the device each
// time.  Here a host-built list of (scene, tile) work items, sorted by cost
// (pulses x tile pixels, largest first), is drained by a persistent grid
// through an atomic work counter, so short and long scenes share the GPU
// without a launch per scene; every item runs the same tile engine as
// sarf_backproject and writes its tile of that scene's image.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sarf_bp_core.hpp"

namespace sarf {

int bp_pick_tile(const sarf_bp_desc *d, double dx, double dy);
int bp_pick_xl_tile(const sarf_bp_desc *d);
int bp_tile_shape(int id, int *tx, int *ty);
int bp_fill_args(sarf_ctx *ctx, const sarf_bp_desc *d, int tile, BpArgs *a);
bool bp_tile_accurate(const sarf_bp_desc *d, int id);

namespace {

struct SceneDev {
    BpArgs a;
    const float2 *prof;
    const double *ant, *r0, *f0;
    float2 *img;
    int shape;  // mixed batches: index of the scene's tile in kMixedIds
};

// Tiles of a mixed batch (all 2048 pixels, lanes along x) and their engines.
constexpr int kMixedIds[3] = {15, 16, 13};
using MixedT0 = BpTile<256, 8, true>;
using MixedT1 = BpTile<128, 8, true>;
using MixedT2 = BpTile<64, 8, true>;
// One engine's LDS (staging and per-column terms); a mixed kernel holds
// the union, no more than the largest engine alone.
template <class Tile>
struct EngineLds {
    typename Tile::Lds L;
    alignas(16) float fx[Tile::kFx];
};
union MixedLds {
    EngineLds<MixedT0> e0;
    EngineLds<MixedT1> e1;
    EngineLds<MixedT2> e2;
};
// the three engines share one occupancy target
static_assert(bp_waves_per_eu<256, 8, true>() == bp_waves_per_eu<128, 8, true>() &&
                  bp_waves_per_eu<128, 8, true>() == bp_waves_per_eu<64, 8, true>(),
              "mixed tiles: one waves-per-EU target");

__global__ void reset_counter(unsigned int *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

template <int TY, int PX, bool XL>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<TY, PX, XL>())))
void bp_scenes(const SceneDev *__restrict__ scenes,
                                                        const int2 *__restrict__ items,
                                                        int nitems,
                                                        unsigned int *__restrict__ counter) {
    __shared__ typename BpTile<TY, PX, XL>::Lds L;
    __shared__ __align__(16) float fx[BpTile<TY, PX, XL>::kFx];
    __shared__ int sItem;
    for (;;) {
        __syncthreads();  // previous item done with sItem and L
        if (threadIdx.x == 0) sItem = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = sItem;
        if (it >= nitems) break;  // every workgroup reaches this once the list is drained
        const int2 w = items[it];
        const SceneDev &S = scenes[w.x];
        const BpArgs a = S.a;
        BpTile<TY, PX, XL> T;
        T.init(a, w.y);
        T.run(a, L, fx, S.prof, PlainPos{S.ant}, S.r0, S.f0, 0, a.Np);
#pragma unroll
        for (int j = 0; j < PX; j++) {
            int x, y;
            if (T.pixel(j, x, y, a)) S.img[(size_t)x * a.Ny + y] = make_float2(T.accr[j], T.acci[j]);
        }
    }
}

// One work item of a scene: tile `tile` of the scene's image with tile
// engine Tile.  Not inlined: each engine keeps its own register allocation
// (inlined side by side, the three needed 170 VGPRs, half the occupancy).
template <class Tile>
__device__ __attribute__((noinline)) void scene_item(const SceneDev &S, int tile,
                                                     typename Tile::Lds &L, float *fx) {
    const BpArgs a = S.a;
    Tile T;
    T.init(a, tile);
    T.run(a, L, fx, S.prof, PlainPos{S.ant}, S.r0, S.f0, 0, a.Np);
#pragma unroll
    for (int j = 0; j < Tile::kPx; j++) {
        int x, y;
        if (T.pixel(j, x, y, a)) S.img[(size_t)x * a.Ny + y] = make_float2(T.accr[j], T.acci[j]);
    }
}

// bp_scenes with a tile shape per scene: an item runs its scene's engine
// (uniform per workgroup); LDS is the union of the engines' own.
__global__ __launch_bounds__(kBpThreads)
__attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<256, 8, true>()))) void bp_scenes_mixed(const SceneDev *__restrict__ scenes, const int2 *__restrict__ items,
                     int nitems, unsigned int *__restrict__ counter) {
    __shared__ MixedLds L;
    __shared__ int sItem;
    for (;;) {
        __syncthreads();  // previous item done with sItem and L
        if (threadIdx.x == 0) sItem = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = sItem;
        if (it >= nitems) break;  // every workgroup reaches this once the list is drained
        const int2 w = items[it];
        const SceneDev &S = scenes[w.x];
        switch (S.shape) {
            case 0: scene_item<MixedT0>(S, w.y, L.e0.L, L.e0.fx); break;
            case 1: scene_item<MixedT1>(S, w.y, L.e1.L, L.e1.fx); break;
            default: scene_item<MixedT2>(S, w.y, L.e2.L, L.e2.fx); break;
        }
    }
}

template <int TY, int PX, bool XL = false>
void launch_scenes_t(int grid, const SceneDev *sc, const int2 *items, int nitems,
                     unsigned int *counter, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, s, counter);
    hipLaunchKernelGGL((bp_scenes<TY, PX, XL>), dim3(grid), dim3(kBpThreads), 0, s, sc, items,
                       nitems, counter);
}

// Tile ids compiled for batches (at most 8 pixels per lane).
constexpr int kSceneTiles[] = {1, 2, 3, 4, 5, 11, 12, 13, 14, 15, 16};
bool scene_tile_ok(int id) {
    for (int t : kSceneTiles)
        if (t == id) return true;
    return false;
}

}  // namespace
}  // namespace sarf

using namespace sarf;

namespace sarf {
namespace {

// Default batch tile: the choice for the finest-spaced scene (it satisfies
// the precision bound for all), at most 8 pixels per lane.
int scenes_default_tile(int nscenes, const sarf_scene *scenes) {
    int tid = 1;
    double best_sp = 1e300;
    for (int i = 0; i < nscenes; i++) {
        const sarf_scene &c = scenes[i];
        if (c.Nx < 2 || c.Ny < 2) continue;
        const double sp = std::max(c.Wx / (c.Nx - 1), c.Wy / (c.Ny - 1));
        if (sp < best_sp) {
            best_sp = sp;
            sarf_bp_desc d{};
            d.Nx = c.Nx; d.Ny = c.Ny; d.Nfft = c.Nfft; d.Np = c.Np;
            d.x0 = c.x0; d.y0 = c.y0; d.Wx = c.Wx; d.Wy = c.Wy; d.maxWr = c.maxWr;
            d.los_x = c.los_x; d.los_y = c.los_y;
            tid = bp_pick_tile(&d, c.Wx / (c.Nx - 1), c.Wy / (c.Ny - 1));
        }
    }
    if (tid == 6) tid = 3;
    if (tid == 7) tid = 4;
    return tid;
}

// True if every scene can take a tile of its own among 15, 16 and 13 and
// has a line-of-sight hint along y (|los_y| >= |los_x|), the orientation
// those lanes-along-x tiles are fast for.
bool scenes_mixed_ok(int nscenes, const sarf_scene *scenes, bool need_hint) {
    for (int i = 0; i < nscenes; i++) {
        const sarf_scene &c = scenes[i];
        if (need_hint && !(c.los_x != 0.0 || c.los_y != 0.0)) return false;
        if (need_hint && std::fabs(c.los_y) < std::fabs(c.los_x)) return false;
        sarf_bp_desc d{};
        d.Nx = c.Nx; d.Ny = c.Ny; d.Nfft = c.Nfft; d.Np = c.Np;
        d.x0 = c.x0; d.y0 = c.y0; d.Wx = c.Wx; d.Wy = c.Wy; d.maxWr = c.maxWr;
        d.los_x = c.los_x; d.los_y = c.los_y;
        if (bp_pick_xl_tile(&d) == 0) return false;
    }
    return true;
}

}  // namespace
}  // namespace sarf

extern "C" int sarf_bp_scenes(sarf_ctx *ctx, int nscenes, const sarf_scene *scenes, int tile,
                              void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (nscenes < 1 || !scenes) return fail(SARF_ERR_INVALID, "sarf_bp_scenes: no scenes");
    std::vector<SceneDev> host(nscenes);
    struct Item {
        int scene, tile;
        double cost;
    };
    std::vector<Item> items;
    // one tile shape for the whole batch (the given one, else the default),
    // or with SARF_SCENES_TILE_MIXED each scene's own among kMixedIds
    // default: a tile per scene when every scene suits it, else one tile
    const bool mixed = tile == SARF_SCENES_TILE_MIXED ||
                       (tile == 0 && scenes_mixed_ok(nscenes, scenes, true));
    const int tid = mixed ? SARF_SCENES_TILE_MIXED
                          : tile > 0 ? tile : scenes_default_tile(nscenes, scenes);
    if (!mixed) {
        int tx = 0, ty = 0;
        if (!bp_tile_shape(tid, &tx, &ty) || !scene_tile_ok(tid))
            return fail(SARF_ERR_INVALID, "sarf_bp_scenes: tile %d not available for batches",
                        tid);
    }
    for (int i = 0; i < nscenes; i++) {
        const sarf_scene &c = scenes[i];
        if (!c.prof || !c.ant || !c.r0 || !c.f0 || !c.img)
            return fail(SARF_ERR_INVALID, "sarf_bp_scenes: scene %d has a null pointer", i);
        sarf_bp_desc d{};
        d.Nx = c.Nx; d.Ny = c.Ny; d.Nfft = c.Nfft; d.Np = c.Np;
        d.x0 = c.x0; d.y0 = c.y0; d.Wx = c.Wx; d.Wy = c.Wy; d.maxWr = c.maxWr;
        d.los_x = c.los_x; d.los_y = c.los_y;
        d.precision = SARF_BP_MIXED;
        int st = tid, shape = 0;
        if (mixed) {
            st = bp_pick_xl_tile(&d);
            if (st == 0)
                return fail(SARF_ERR_INVALID, "sarf_bp_scenes: scene %d has no accurate tile "
                            "among 15, 16 and 13 for a mixed batch", i);
            while (shape < 3 && kMixedIds[shape] != st) shape++;
            if (shape == 3)
                return fail(SARF_ERR_INVALID, "sarf_bp_scenes: scene %d: tile %d has no engine in "
                            "a mixed batch", i, st);
        }
        int tx = 0, ty = 0;
        bp_tile_shape(st, &tx, &ty);
        BpArgs a{};
        const int rc = bp_fill_args(ctx, &d, st, &a);
        if (rc != SARF_OK) return fail(rc, "sarf_bp_scenes: scene %d: %s", i, sarf_last_error());
        host[i] = SceneDev{a, (const float2 *)c.prof, c.ant, c.r0, c.f0, (float2 *)c.img, shape};
        const int nt = a.tiles_x * a.tiles_y;
        for (int t = 0; t < nt; t++) items.push_back(Item{i, t, (double)c.Np * tx * ty});
    }
    // largest first; within a scene keep tile order (stable sort)
    std::stable_sort(items.begin(), items.end(),
                     [](const Item &p, const Item &q) { return p.cost > q.cost; });
    std::vector<int2> hitems(items.size());
    for (size_t i = 0; i < items.size(); i++) hitems[i] = make_int2(items[i].scene, items[i].tile);
    const size_t bytes_sc = sizeof(SceneDev) * nscenes, bytes_it = sizeof(int2) * hitems.size();
    hipStream_t s = (hipStream_t)stream;
    // The work list depends only on the caller's descriptors and the tile: a
    // list seen before has its own device copy, so the call launches without
    // copying or waiting (and can be captured into a graph once it has run
    // outside the capture).
    std::vector<unsigned char> key(sizeof(int) + sizeof(sarf_scene) * (size_t)nscenes);
    std::memcpy(key.data(), &tid, sizeof(int));
    std::memcpy(key.data() + sizeof(int), scenes, sizeof(sarf_scene) * (size_t)nscenes);
    void *buf = nullptr;
    bool fresh = false;
    SARF_TRY(list_buffer(ctx, 's', key, bytes_sc + bytes_it + 256, s, "sarf_bp_scenes", &buf,
                         &fresh));
    SceneDev *dsc = (SceneDev *)buf;
    int2 *ditems = (int2 *)(((uintptr_t)((char *)buf + bytes_sc) + 15) & ~(uintptr_t)15);
    unsigned int *counter = (unsigned int *)(((uintptr_t)((char *)ditems + bytes_it) + 15) & ~(uintptr_t)15);
    if (fresh) {
        // pageable source: the copy completes before returning, so the host
        // vectors may go out of scope
        SARF_HIP(hipMemcpyAsync(dsc, host.data(), bytes_sc, hipMemcpyHostToDevice, s));
        SARF_HIP(hipMemcpyAsync(ditems, hitems.data(), bytes_it, hipMemcpyHostToDevice, s));
        SARF_HIP(hipStreamSynchronize(s));
    }
    const int cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    const int nit = (int)hitems.size();
    const int grid = std::min(nit, 4 * cus);
    switch (tid) {
        case SARF_SCENES_TILE_MIXED:
            hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, s, counter);
            hipLaunchKernelGGL(bp_scenes_mixed, dim3(grid), dim3(kBpThreads), 0, s, dsc, ditems,
                               nit, counter);
            break;
        case 1: launch_scenes_t<32, 4>(grid, dsc, ditems, nit, counter, s); break;
        case 2: launch_scenes_t<64, 4>(grid, dsc, ditems, nit, counter, s); break;
        case 3: launch_scenes_t<64, 8>(grid, dsc, ditems, nit, counter, s); break;
        case 4: launch_scenes_t<32, 8>(grid, dsc, ditems, nit, counter, s); break;
        case 5: launch_scenes_t<16, 4>(grid, dsc, ditems, nit, counter, s); break;
        case 11: launch_scenes_t<32, 4, true>(grid, dsc, ditems, nit, counter, s); break;
        case 12: launch_scenes_t<64, 4, true>(grid, dsc, ditems, nit, counter, s); break;
        case 13: launch_scenes_t<64, 8, true>(grid, dsc, ditems, nit, counter, s); break;
        case 14: launch_scenes_t<32, 8, true>(grid, dsc, ditems, nit, counter, s); break;
        case 15: launch_scenes_t<256, 8, true>(grid, dsc, ditems, nit, counter, s); break;
        case 16: launch_scenes_t<128, 8, true>(grid, dsc, ditems, nit, counter, s); break;
    }
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

// Times sarf_bp_scenes with the default batch tile, every other batch tile
// that meets the accuracy bound for all scenes and, if every scene has an
// accurate one of its tiles, SARF_SCENES_TILE_MIXED, and returns the fastest
// (1 % hysteresis towards the earlier candidate, in that order); waits for
// the stream.  The images hold the last run's result.
extern "C" int sarf_autotune_scenes(sarf_ctx *ctx, int nscenes, const sarf_scene *scenes, int reps,
                                    int *best_tile, float *best_ms, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (nscenes < 1 || !scenes || !best_tile || reps < 1 || reps > 100)
        return fail(SARF_ERR_INVALID, "sarf_autotune_scenes: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SARF_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(SARF_ERR_HIP, "sarf_autotune_scenes: event creation failed");
    }
    const int def = scenes_mixed_ok(nscenes, scenes, true) ? SARF_SCENES_TILE_MIXED
                                                            : scenes_default_tile(nscenes, scenes);
    std::vector<int> cand{def};
    for (int id : kSceneTiles) {
        if (id == def) continue;
        bool ok = true;
        for (int i = 0; i < nscenes && ok; i++) {
            const sarf_scene &c = scenes[i];
            sarf_bp_desc d{};
            d.Nx = c.Nx; d.Ny = c.Ny; d.Nfft = c.Nfft; d.Np = c.Np;
            d.x0 = c.x0; d.y0 = c.y0; d.Wx = c.Wx; d.Wy = c.Wy; d.maxWr = c.maxWr;
            d.los_x = c.los_x; d.los_y = c.los_y;
            ok = bp_tile_accurate(&d, id);
        }
        if (ok) cand.push_back(id);
    }
    // a tile per scene, if every scene has one
    if (def != SARF_SCENES_TILE_MIXED && scenes_mixed_ok(nscenes, scenes, false))
        cand.push_back(SARF_SCENES_TILE_MIXED);
    int best = 0;
    float bms = 0.f;
    int rc = SARF_OK;
    for (int id : cand) {
        rc = sarf_bp_scenes(ctx, nscenes, scenes, id, stream);  // upload + warm
        if (rc != SARF_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(SARF_ERR_HIP, "event record"); break; }
        for (int r = 0; r < reps && rc == SARF_OK; r++) rc = sarf_bp_scenes(ctx, nscenes, scenes, id, stream);
        if (rc != SARF_OK) break;
        float ms = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(SARF_ERR_HIP, "sarf_autotune_scenes: timing failed");
            break;
        }
        ms /= reps;
        if (best == 0 || ms < 0.99f * bms) {
            best = id;
            bms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SARF_OK) return rc;
    *best_tile = best;
    if (best_ms) *best_ms = bms;
    return SARF_OK;
}
