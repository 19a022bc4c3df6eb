// Batched atmosphere LUT builds for gfx950: many atmosphere sets, each with its
// own transmittance and multiple-scattering LUTs, and many (set, sun) views,
// each with its own sky-view LUT, at mixed resolutions, built by one hipGraph
// launch.
//
// The graph holds four kernel nodes: a counter reset, then one persistent
// kernel per pass (transmittance -> multiple scattering -> sky view; kernel
// order carries the dependencies).  Each persistent kernel runs a fixed number
// of workgroups per CU that pull tiles (item, first texel) from a device
// counter until the tile list is exhausted, so items of very different sizes
// share the machine without one launch per item.  Every texel runs the
// single-set arithmetic of lut_texels.h.  Tile sizes, workgroup sizes,
// workgroups per CU and the sky-view tile order are chosen by the autotuner
// below, which times each pass's candidates on the batch itself.
#include <algorithm>
#include <vector>

#include "lut_texels.h"

namespace skyhip {

int validate_atmosphere(const SkyhipAtmosphereBufferData* b);
int validate_render(const SkyhipAtmosphereRenderBufferData* b);

struct BatchSetDev {
    int tw, th, mw, mh;
    long long t_off, ms_off;  // float4 offsets into the pools
};
struct BatchViewDev {
    int set, sw, sh, pad;
    long long sv_off;
};
struct Tile {
    uint32_t item, first;
};

enum { kCtrT = 0, kCtrMS = 1, kCtrSV = 2 };

__global__ void batch_reset_kernel(unsigned* __restrict__ ctr) {
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0u;
}

// Pulls the next tile index for the whole workgroup.
__device__ __forceinline__ int next_tile(unsigned* ctr, int* s_tile) {
    if (threadIdx.x == 0) *s_tile = int(atomicAdd(ctr, 1u));
    __syncthreads();
    int tile = *s_tile;
    __syncthreads();
    return tile;
}

// Cell-corner copies of every item's T (kind 0) or MS (kind 1) LUT: texel (x, y) of an item is
// followed by the rest of its bilinear footprint, (x, y), (x1, y), (x, y1), (x1, y1) with
// x1 = min(x + 1, w - 1) and y1 likewise (4 float4 per texel, lut_bilinear<true>).  Items stride
// over blockIdx.y, texels over blockIdx.x.
__global__ __launch_bounds__(256) void batch_lut_cells_kernel(const BatchSetDev* __restrict__ sets, int n_sets,
                                                              int kind, const float4* __restrict__ pool,
                                                              float4* __restrict__ cells) {
    for (int item = blockIdx.y; item < n_sets; item += gridDim.y) {
        const BatchSetDev S = sets[item];
        const int w = kind == 0 ? S.tw : S.mw, h = kind == 0 ? S.th : S.mh;
        const long long off = kind == 0 ? S.t_off : S.ms_off;
        const float4* src = pool + off;
        for (int t = blockIdx.x * 256 + threadIdx.x; t < w * h; t += gridDim.x * 256) {
            const int x = t % w, y = t / w, x1 = min(x + 1, w - 1), y1 = min(y + 1, h - 1);
            float4* dst = cells + 4 * (off + t);
            dst[0] = src[y * w + x];
            dst[1] = src[y * w + x1];
            dst[2] = src[y1 * w + x];
            dst[3] = src[y1 * w + x1];
        }
    }
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void batch_transmittance_kernel(
    const SkyhipAtmosphereBufferData* __restrict__ atm, const BatchSetDev* __restrict__ sets,
    const Tile* __restrict__ tiles, int n_tiles, int tile_size, float4* __restrict__ pool, unsigned* ctr) {
    __shared__ int s_tile;
    for (;;) {
        const int tile = next_tile(ctr + kCtrT, &s_tile);
        if (tile >= n_tiles) return;  // workgroup-uniform
        const Tile tl = tiles[tile];
        const BatchSetDev S = sets[tl.item];
        const SkyhipAtmosphereBufferData b = atm[tl.item];
        const int n = S.tw * S.th;
        for (int k = threadIdx.x; k < tile_size; k += kBlock) {
            const int t = int(tl.first) + k;
            if (t < n) pool[S.t_off + t] = transmittance_texel_rgba(b, t % S.tw, t / S.tw, S.tw, S.th);
        }
    }
}

template <int kWaves>
__global__ __launch_bounds__(64 * kWaves) void batch_multiscattering_kernel(
    const SkyhipAtmosphereBufferData* __restrict__ atm, const BatchSetDev* __restrict__ sets,
    const Tile* __restrict__ tiles, int n_tiles, const float4* __restrict__ t_cells, float4* __restrict__ pool,
    unsigned* ctr) {
    __shared__ int s_tile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (;;) {
        const int tile = next_tile(ctr + kCtrMS, &s_tile);
        if (tile >= n_tiles) return;
        const Tile tl = tiles[tile];
        const BatchSetDev S = sets[tl.item];
        const int texel = int(tl.first) + wave;
        if (texel < S.mw * S.mh) {  // wave-uniform
            const AtmDev A = make_atm(atm[tl.item]);
            float4 v = multiscattering_texel<true>(A, t_cells + 4 * S.t_off, S.tw, S.th, texel % S.mw, texel / S.mw,
                                                   S.mw, S.mh, lane);
            if (lane == 0) pool[S.ms_off + texel] = v;
        }
    }
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void batch_sky_view_kernel(
    const SkyhipAtmosphereBufferData* __restrict__ atm, const SkyhipAtmosphereRenderBufferData* __restrict__ renders,
    const BatchSetDev* __restrict__ sets, const BatchViewDev* __restrict__ views, const Tile* __restrict__ tiles,
    int n_tiles, int tile_size, const float4* __restrict__ t_cells, const float4* __restrict__ ms_cells,
    float4* __restrict__ l_pool, float4* __restrict__ tr_pool, unsigned* ctr) {
    __shared__ int s_tile;
    for (;;) {
        const int tile = next_tile(ctr + kCtrSV, &s_tile);
        if (tile >= n_tiles) return;
        const Tile tl = tiles[tile];
        const BatchViewDev V = views[tl.item];
        const BatchSetDev S = sets[V.set];
        const AtmDev A = make_atm(atm[V.set]);
        const RenderDev R = make_render(renders[tl.item]);
        const int n = V.sw * V.sh;
        for (int k = threadIdx.x; k < tile_size; k += kBlock) {
            const int t = int(tl.first) + k;
            if (t >= n) break;
            F3 L, Tr;
            sky_view_texel<true>(A, R, t_cells + 4 * S.t_off, S.tw, S.th, ms_cells + 4 * S.ms_off, S.mw, S.mh,
                                 nullptr, 1, t % V.sw, t / V.sw, V.sw, V.sh, L, Tr);
            l_pool[V.sv_off + t] = make_float4(L.x, L.y, L.z, 0.0f);
            tr_pool[V.sv_off + t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
        }
    }
}

static int loop_iterations(float steps) {  // for (float i = 0.5; i < steps; ++i)
    int n = 0;
    for (float i = 0.5f; i < steps && n < (1 << 20); i += 1.0f) ++n;
    return n;
}

}  // namespace skyhip

using namespace skyhip;

struct SkyhipLutBatch {
    SkyhipContext* ctx = nullptr;
    int n_sets = 0, n_views = 0;
    std::vector<BatchSetDev> sets;
    std::vector<BatchViewDev> views;
    SkyhipAtmosphereBufferData* d_atm = nullptr;
    SkyhipAtmosphereRenderBufferData* d_render = nullptr;
    BatchSetDev* d_sets = nullptr;
    BatchViewDev* d_views = nullptr;
    float4 *t_pool = nullptr, *ms_pool = nullptr, *l_pool = nullptr, *tr_pool = nullptr;
    float4 *t_cells = nullptr, *ms_cells = nullptr;  // cell-corner copies of the T and MS pools (4x)
    int max_t_texels = 0, max_ms_texels = 0;
    unsigned* d_ctr = nullptr;
    Tile *d_t_tiles = nullptr, *d_ms_tiles = nullptr, *d_sv_tiles = nullptr;
    int n_t_tiles = 0, n_ms_tiles = 0, n_sv_tiles = 0;
    SkyhipLutBatchTuning tuning{};
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long steps = 0;
};

namespace {

void release_graph(SkyhipLutBatch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_tiles(SkyhipLutBatch* b) {
    for (Tile** p : {&b->d_t_tiles, &b->d_ms_tiles, &b->d_sv_tiles}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
}

int check_tuning(const SkyhipLutBatchTuning& t) {
    auto block_ok = [](int v) { return v == 64 || v == 128 || v == 256; };
    auto waves_ok = [](int v) { return v == 1 || v == 2 || v == 4 || v == 8; };
    if (!block_ok(t.transmittance_block) || !block_ok(t.sky_view_block) || !waves_ok(t.multiscattering_waves))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "block sizes must be 64/128/256 and waves 1/2/4/8");
    if (t.transmittance_tile < t.transmittance_block || t.transmittance_tile > 65536 ||
        t.sky_view_tile < t.sky_view_block || t.sky_view_tile > 65536)
        return fail(SKYHIP_ERR_INVALID_SIZE, "tile sizes must be in [block, 65536]");
    for (int v : {t.transmittance_blocks_per_cu, t.multiscattering_blocks_per_cu, t.sky_view_blocks_per_cu})
        if (v < 1 || v > 32) return fail(SKYHIP_ERR_INVALID_SIZE, "workgroups per CU must be in [1, 32]");
    return SKYHIP_OK;
}

int upload_tiles(const std::vector<Tile>& h, Tile** d) {
    SKYHIP_HIP(hipMalloc(reinterpret_cast<void**>(d), std::max<size_t>(1, h.size()) * sizeof(Tile)));
    if (!h.empty()) SKYHIP_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(Tile), hipMemcpyHostToDevice));
    return SKYHIP_OK;
}

// Tile lists for the current tuning.
int build_tiles(SkyhipLutBatch* b) {
    free_tiles(b);
    const SkyhipLutBatchTuning& t = b->tuning;
    std::vector<Tile> tt, mt, st;
    for (int s = 0; s < b->n_sets; ++s) {
        const BatchSetDev& S = b->sets[s];
        for (int f = 0; f < S.tw * S.th; f += t.transmittance_tile) tt.push_back({uint32_t(s), uint32_t(f)});
        for (int f = 0; f < S.mw * S.mh; f += t.multiscattering_waves) mt.push_back({uint32_t(s), uint32_t(f)});
    }
    std::vector<int> order(b->n_views);
    for (int v = 0; v < b->n_views; ++v) order[v] = v;
    if (t.sky_view_largest_first)
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return b->views[x].sw * b->views[x].sh > b->views[y].sw * b->views[y].sh;
        });
    for (int v : order) {
        const BatchViewDev& V = b->views[v];
        for (int f = 0; f < V.sw * V.sh; f += t.sky_view_tile) st.push_back({uint32_t(v), uint32_t(f)});
    }
    b->n_t_tiles = int(tt.size());
    b->n_ms_tiles = int(mt.size());
    b->n_sv_tiles = int(st.size());
    SKYHIP_CHECK(upload_tiles(tt, &b->d_t_tiles));
    SKYHIP_CHECK(upload_tiles(mt, &b->d_ms_tiles));
    SKYHIP_CHECK(upload_tiles(st, &b->d_sv_tiles));
    return SKYHIP_OK;
}

unsigned grid_for(const SkyhipLutBatch* b, int per_cu, int n_tiles) {
    return unsigned(std::max(1, std::min(b->ctx->cu_count * per_cu, std::max(n_tiles, 1))));
}

dim3 cells_grid(int max_texels, int n_sets) {
    return dim3(std::min(div_up(max_texels, 256), 4096u), unsigned(std::min(n_sets, 32768)));
}

void enqueue_t(SkyhipLutBatch* b, hipStream_t st) {
    const SkyhipLutBatchTuning& t = b->tuning;
    dim3 g(grid_for(b, t.transmittance_blocks_per_cu, b->n_t_tiles));
#define SKYHIP_T(B)                                                                                              \
    hipLaunchKernelGGL(batch_transmittance_kernel<B>, g, dim3(B), 0, st, b->d_atm, b->d_sets, b->d_t_tiles,    \
                       b->n_t_tiles, t.transmittance_tile, b->t_pool, b->d_ctr)
    if (t.transmittance_block == 64) SKYHIP_T(64);
    else if (t.transmittance_block == 128) SKYHIP_T(128);
    else SKYHIP_T(256);
#undef SKYHIP_T
}

// The multiple-scattering pass, with the transmittance cell copy it reads built first.
void enqueue_ms(SkyhipLutBatch* b, hipStream_t st) {
    const SkyhipLutBatchTuning& t = b->tuning;
    hipLaunchKernelGGL(batch_lut_cells_kernel, cells_grid(b->max_t_texels, b->n_sets), dim3(256), 0, st, b->d_sets,
                       b->n_sets, 0, b->t_pool, b->t_cells);
    dim3 g(grid_for(b, t.multiscattering_blocks_per_cu, b->n_ms_tiles));
#define SKYHIP_MS(W)                                                                                              \
    hipLaunchKernelGGL(batch_multiscattering_kernel<W>, g, dim3(64 * W), 0, st, b->d_atm, b->d_sets,            \
                       b->d_ms_tiles, b->n_ms_tiles, b->t_cells, b->ms_pool, b->d_ctr)
    if (t.multiscattering_waves == 1) SKYHIP_MS(1);
    else if (t.multiscattering_waves == 2) SKYHIP_MS(2);
    else if (t.multiscattering_waves == 4) SKYHIP_MS(4);
    else SKYHIP_MS(8);
#undef SKYHIP_MS
}

// The sky-view pass, with the multiple-scattering cell copy it reads built first (the
// transmittance one comes from the multiple-scattering pass).
void enqueue_sv(SkyhipLutBatch* b, hipStream_t st) {
    const SkyhipLutBatchTuning& t = b->tuning;
    hipLaunchKernelGGL(batch_lut_cells_kernel, cells_grid(b->max_ms_texels, b->n_sets), dim3(256), 0, st, b->d_sets,
                       b->n_sets, 1, b->ms_pool, b->ms_cells);
    dim3 g(grid_for(b, t.sky_view_blocks_per_cu, b->n_sv_tiles));
#define SKYHIP_SV(B)                                                                                              \
    hipLaunchKernelGGL(batch_sky_view_kernel<B>, g, dim3(B), 0, st, b->d_atm, b->d_render, b->d_sets, b->d_views, \
                       b->d_sv_tiles, b->n_sv_tiles, t.sky_view_tile, b->t_cells, b->ms_cells, b->l_pool,        \
                       b->tr_pool, b->d_ctr)
    if (t.sky_view_block == 64) SKYHIP_SV(64);
    else if (t.sky_view_block == 128) SKYHIP_SV(128);
    else SKYHIP_SV(256);
#undef SKYHIP_SV
}

// Captures reset -> T -> MS -> SV into the batch's graph.
int capture(SkyhipLutBatch* b) {
    release_graph(b);
    SKYHIP_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeRelaxed));
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, b->capture, b->d_ctr);
    enqueue_t(b, b->capture);
    enqueue_ms(b, b->capture);
    enqueue_sv(b, b->capture);
    hipError_t e = hipGetLastError();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->capture, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return check_hip(e != hipSuccess ? e : e2, "batch graph capture");
    }
    b->graph = g;
    SKYHIP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return SKYHIP_OK;
}

// Rebuilds the tile lists and the graph for tuning t.  The graph is released before the tile
// lists it points to are freed, so a failure part-way leaves no graph that reads freed memory:
// the previous tuning is then restored, or, if that fails too, the batch is left without a
// graph (skyhip_lut_batch_launch reports it) rather than with a dangling one.
int apply_tuning(SkyhipLutBatch* b, const SkyhipLutBatchTuning& t) {
    SKYHIP_CHECK(check_tuning(t));
    SKYHIP_HIP(hipStreamSynchronize(b->ctx->stream));
    const SkyhipLutBatchTuning old = b->tuning;
    release_graph(b);
    b->tuning = t;
    int rc = build_tiles(b);
    if (rc == SKYHIP_OK) rc = capture(b);
    if (rc == SKYHIP_OK) return SKYHIP_OK;
    release_graph(b);
    free_tiles(b);
    if (std::memcmp(&old, &t, sizeof(t)) != 0 && check_tuning(old) == SKYHIP_OK) {
        b->tuning = old;
        if (build_tiles(b) != SKYHIP_OK || capture(b) != SKYHIP_OK) {
            release_graph(b);
            free_tiles(b);
        }
    }
    return rc;
}

// Median time (ms) of `reps` runs of one pass (with its counter reset) on the context stream.
template <class F>
int time_pass(SkyhipLutBatch* b, F enqueue, int reps, float* out_ms) {
    hipStream_t st = b->ctx->stream;
    std::vector<float> ms;
    for (int r = 0; r < reps + 1; ++r) {
        hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, b->d_ctr);
        SKYHIP_HIP(hipEventRecord(b->ctx->events[60], st));
        enqueue(b, st);
        SKYHIP_HIP(hipEventRecord(b->ctx->events[61], st));
        SKYHIP_HIP(hipEventSynchronize(b->ctx->events[61]));
        SKYHIP_LAUNCH_CHECK("batch pass");
        float v = 0.0f;
        SKYHIP_HIP(hipEventElapsedTime(&v, b->ctx->events[60], b->ctx->events[61]));
        if (r > 0) ms.push_back(v);  // the first run warms up
    }
    std::sort(ms.begin(), ms.end());
    *out_ms = ms[ms.size() / 2];
    return SKYHIP_OK;
}

SkyhipLutBatchTuning default_tuning() {
    SkyhipLutBatchTuning t{};
    t.transmittance_block = 256;
    t.transmittance_tile = 1024;
    t.transmittance_blocks_per_cu = 4;
    t.multiscattering_waves = 4;
    t.multiscattering_blocks_per_cu = 8;
    t.sky_view_block = 256;
    t.sky_view_tile = 1024;
    t.sky_view_blocks_per_cu = 4;
    t.sky_view_largest_first = 1;
    return t;
}

}  // namespace

extern "C" int skyhip_lut_batch_create(SkyhipContext* ctx, const SkyhipLutBatchSet* sets, int32_t set_count,
                                       const SkyhipLutBatchView* views, int32_t view_count, SkyhipLutBatch** out) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!sets || !views || !out) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (set_count < 1 || set_count > 65536 || view_count < 0 || view_count > (1 << 20))
        return fail(SKYHIP_ERR_INVALID_SIZE, "set count outside [1, 65536] or view count outside [0, 2^20]");
    SkyhipLutBatch* b = new SkyhipLutBatch();
    b->ctx = ctx;
    b->n_sets = set_count;
    b->n_views = view_count;
    auto bad = [&](int rc) {
        skyhip_lut_batch_destroy(b);
        return rc;
    };
    long long t_total = 0, ms_total = 0, sv_total = 0;
    std::vector<SkyhipAtmosphereBufferData> atm(set_count);
    for (int s = 0; s < set_count; ++s) {
        const SkyhipLutBatchSet& S = sets[s];
        int rc = validate_atmosphere(&S.atmosphere);
        if (rc != SKYHIP_OK) return bad(rc);
        for (int v : {S.transmittance_width, S.transmittance_height, S.multiscattering_width,
                      S.multiscattering_height})
            if (v < 2 || v > 4096) return bad(fail(SKYHIP_ERR_INVALID_SIZE, "set %d: LUT size outside [2, 4096]", s));
        atm[s] = S.atmosphere;
        BatchSetDev d{S.transmittance_width, S.transmittance_height, S.multiscattering_width,
                      S.multiscattering_height, t_total, ms_total};
        b->sets.push_back(d);
        t_total += (long long)d.tw * d.th;
        ms_total += (long long)d.mw * d.mh;
        b->max_t_texels = std::max(b->max_t_texels, d.tw * d.th);
        b->max_ms_texels = std::max(b->max_ms_texels, d.mw * d.mh);
        b->steps += (unsigned long long)d.tw * d.th * loop_iterations(S.atmosphere.transmittance_steps) +
                    64ull * d.mw * d.mh * loop_iterations(S.atmosphere.multiscattering_steps);
    }
    std::vector<SkyhipAtmosphereRenderBufferData> rnd(view_count);
    for (int v = 0; v < view_count; ++v) {
        const SkyhipLutBatchView& V = views[v];
        if (V.set < 0 || V.set >= set_count) return bad(fail(SKYHIP_ERR_INVALID_ARGUMENT, "view %d: bad set", v));
        int rc = validate_render(&V.render);
        if (rc != SKYHIP_OK) return bad(rc);
        if (V.sky_view_width < 2 || V.sky_view_height < 2 || V.sky_view_width > 4096 || V.sky_view_height > 4096)
            return bad(fail(SKYHIP_ERR_INVALID_SIZE, "view %d: sky-view size outside [2, 4096]", v));
        rnd[v] = V.render;
        b->views.push_back({V.set, V.sky_view_width, V.sky_view_height, 0, sv_total});
        sv_total += (long long)V.sky_view_width * V.sky_view_height;
        b->steps += (unsigned long long)V.sky_view_width * V.sky_view_height *
                    loop_iterations(V.render.sky_view_lut_steps);
    }
    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
        if (e != hipSuccess) return check_hip(e, "hipMalloc (batch)");
        return SKYHIP_OK;
    };
    int rc = SKYHIP_OK;
    if ((rc = alloc(reinterpret_cast<void**>(&b->d_atm), sizeof(SkyhipAtmosphereBufferData) * set_count)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->d_render), sizeof(SkyhipAtmosphereRenderBufferData) * view_count)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->d_sets), sizeof(BatchSetDev) * set_count)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->d_views), sizeof(BatchViewDev) * view_count)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->t_pool), sizeof(float4) * t_total)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->ms_pool), sizeof(float4) * ms_total)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->t_cells), 4 * sizeof(float4) * t_total)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->ms_cells), 4 * sizeof(float4) * ms_total)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->l_pool), sizeof(float4) * sv_total)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->tr_pool), sizeof(float4) * sv_total)) ||
        (rc = alloc(reinterpret_cast<void**>(&b->d_ctr), 16)))
        return bad(rc);
    auto up = [&](void* d, const void* h, size_t bytes) -> int {
        if (bytes == 0) return SKYHIP_OK;
        hipError_t e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
        return e == hipSuccess ? SKYHIP_OK : check_hip(e, "hipMemcpy (batch)");
    };
    if ((rc = up(b->d_atm, atm.data(), sizeof(SkyhipAtmosphereBufferData) * set_count)) ||
        (rc = up(b->d_render, rnd.data(), sizeof(SkyhipAtmosphereRenderBufferData) * view_count)) ||
        (rc = up(b->d_sets, b->sets.data(), sizeof(BatchSetDev) * set_count)) ||
        (rc = up(b->d_views, b->views.data(), sizeof(BatchViewDev) * view_count)))
        return bad(rc);
    hipError_t e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking);
    if (e != hipSuccess) return bad(check_hip(e, "hipStreamCreate (batch)"));
    if ((rc = apply_tuning(b, default_tuning())) != SKYHIP_OK) return bad(rc);
    *out = b;
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_destroy(SkyhipLutBatch* b) {
    if (!b) return SKYHIP_OK;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    release_graph(b);
    free_tiles(b);
    for (void* p : {(void*)b->d_atm, (void*)b->d_render, (void*)b->d_sets, (void*)b->d_views, (void*)b->t_pool,
                    (void*)b->ms_pool, (void*)b->t_cells, (void*)b->ms_cells, (void*)b->l_pool, (void*)b->tr_pool,
                    (void*)b->d_ctr})
        if (p) (void)hipFree(p);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    delete b;
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_launch(SkyhipLutBatch* b) {
    if (!b || !b->exec) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "batch is NULL or has no graph");
    SKYHIP_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_get_tuning(const SkyhipLutBatch* b, SkyhipLutBatchTuning* out) {
    if (!b || !out) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = b->tuning;
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_set_tuning(SkyhipLutBatch* b, const SkyhipLutBatchTuning* t) {
    if (!b || !t) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    return apply_tuning(b, *t);
}

extern "C" int skyhip_lut_batch_autotune(SkyhipLutBatch* b, int32_t repetitions, SkyhipLutBatchTuning* out_best,
                                         float* out_ms) {
    if (!b) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (repetitions < 1 || repetitions > 100) return fail(SKYHIP_ERR_INVALID_SIZE, "repetitions outside [1, 100]");
    // Valid T and MS pools first: the MS and SV candidates read them.
    SKYHIP_CHECK(skyhip_lut_batch_launch(b));
    SKYHIP_HIP(hipStreamSynchronize(b->ctx->stream));
    const SkyhipLutBatchTuning initial = b->tuning;
    SkyhipLutBatchTuning best = b->tuning;
    // The candidates rebuild the tile lists the graph reads: drop the graph first, and on any
    // failure re-apply the tuning the batch had (apply_tuning leaves it launchable or graph-less).
    release_graph(b);
    auto restore = [&](int rc) {
        (void)apply_tuning(b, initial);
        return rc;
    };
    struct Knob {
        int SkyhipLutBatchTuning::*field;
        std::vector<int> values;
        int pass;  // 0 T, 1 MS, 2 SV
    };
    const std::vector<Knob> knobs = {
        {&SkyhipLutBatchTuning::transmittance_block, {64, 128, 256}, 0},
        {&SkyhipLutBatchTuning::transmittance_tile, {256, 512, 1024, 2048, 4096}, 0},
        {&SkyhipLutBatchTuning::transmittance_blocks_per_cu, {1, 2, 4, 8}, 0},
        {&SkyhipLutBatchTuning::multiscattering_waves, {1, 2, 4, 8}, 1},
        {&SkyhipLutBatchTuning::multiscattering_blocks_per_cu, {2, 4, 8, 16}, 1},
        {&SkyhipLutBatchTuning::sky_view_block, {64, 128, 256}, 2},
        {&SkyhipLutBatchTuning::sky_view_tile, {256, 512, 1024, 2048, 4096}, 2},
        {&SkyhipLutBatchTuning::sky_view_blocks_per_cu, {1, 2, 4, 8, 16}, 2},
        {&SkyhipLutBatchTuning::sky_view_largest_first, {0, 1}, 2},
    };
    float pass_ms[3] = {0, 0, 0};
    for (int round = 0; round < 2; ++round) {  // coordinate descent, two sweeps
        for (const Knob& k : knobs) {
            int best_v = best.*(k.field);
            float best_ms = 3.0e38f;
            for (int v : k.values) {
                SkyhipLutBatchTuning cand = best;
                cand.*(k.field) = v;
                if (check_tuning(cand) != SKYHIP_OK) continue;
                b->tuning = cand;
                int rc = build_tiles(b);
                float ms = 0.0f;
                if (rc == SKYHIP_OK && k.pass == 0) rc = time_pass(b, enqueue_t, repetitions, &ms);
                if (rc == SKYHIP_OK && k.pass == 1) rc = time_pass(b, enqueue_ms, repetitions, &ms);
                if (rc == SKYHIP_OK && k.pass == 2) rc = time_pass(b, enqueue_sv, repetitions, &ms);
                if (rc != SKYHIP_OK) return restore(rc);
                if (ms < best_ms) best_ms = ms, best_v = v;
            }
            best.*(k.field) = best_v;
            pass_ms[k.pass] = best_ms;
        }
    }
    b->tuning = initial;  // apply_tuning's fallback if `best` cannot be applied
    const int rc = apply_tuning(b, best);
    if (rc != SKYHIP_OK) return rc;
    if (out_best) *out_best = best;
    if (out_ms) *out_ms = pass_ms[0] + pass_ms[1] + pass_ms[2];
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_set_outputs(SkyhipLutBatch* b, int32_t set, const float** transmittance,
                                            const float** multiscattering) {
    if (!b) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (set < 0 || set >= b->n_sets) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "set %d out of range", set);
    if (transmittance) *transmittance = reinterpret_cast<const float*>(b->t_pool + b->sets[set].t_off);
    if (multiscattering) *multiscattering = reinterpret_cast<const float*>(b->ms_pool + b->sets[set].ms_off);
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_view_outputs(SkyhipLutBatch* b, int32_t view, const float** luminance,
                                             const float** sky_transmittance) {
    if (!b) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (view < 0 || view >= b->n_views) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "view %d out of range", view);
    if (luminance) *luminance = reinterpret_cast<const float*>(b->l_pool + b->views[view].sv_off);
    if (sky_transmittance) *sky_transmittance = reinterpret_cast<const float*>(b->tr_pool + b->views[view].sv_off);
    return SKYHIP_OK;
}

extern "C" int skyhip_lut_batch_integrator_steps(const SkyhipLutBatch* b, unsigned long long* out_steps) {
    if (!b || !out_steps) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out_steps = b->steps;
    return SKYHIP_OK;
}
