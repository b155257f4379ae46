// st_api.hip - C ABI of the tracer (include/soltrace_hip.h): contexts, scenes,
// single and batched traces. The serving loop lives in st_server.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "st_api_internal.h"
#include "st_bvh.h"

using namespace st;

// ---------------------------------------------------------------------------
// errors and fault injection
namespace st {
static thread_local std::string g_last_error;
static std::atomic<int> g_fail_alloc{0}, g_fail_launch{0};

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(ST_ERR_DEVICE, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

bool take_injected(std::atomic<int>& c) {
    int v = c.load();
    while (v > 0)
        if (c.compare_exchange_weak(v, v - 1)) return true;
    return false;
}

int dev_alloc(void** p, size_t bytes, const char* what) {
    *p = nullptr;
    if (take_injected(g_fail_alloc)) return fail(ST_ERR_INJECTED, "injected allocation failure (%s)", what);
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? ST_ERR_ALLOC : ST_ERR_DEVICE, "hipMalloc(%zu) for %s: %s", bytes, what,
                    hipGetErrorString(e));
    }
    return ST_OK;
}

bool launch_injected() { return take_injected(g_fail_launch); }

void DevBuf::reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
int DevBuf::ensure(size_t n, const char* what) {
    if (ptr && bytes >= n) return ST_OK;
    reset();
    int rc = dev_alloc(&ptr, n, what);
    if (rc == ST_OK) bytes = n;
    return rc;
}

LaunchShape pick_shape(uint64_t rays, int32_t block, int32_t grid, int32_t kernel, int32_t bvh_depth,
                       const LaunchShape* tuned, int32_t min_elements, int32_t waves) {
    LaunchShape sh;
    const int cus = device_cu_count();
    // ST_KERNEL_MEGA: whole ray per lane; ST_KERNEL_QUEUED: continuing rays
    // regrouped in full waves; AUTO keeps the tuned (or default) form
    // an explicit grouped form is a queued form that must keep a grouped variant
    const bool want_grouped = kernel == ST_KERNEL_QUEUED_GROUPED;
    const bool want_queued = kernel == ST_KERNEL_QUEUED || want_grouped;
    const bool use_tuned = tuned && block <= 0 && grid <= 0 && waves <= 0 &&
                           (kernel == ST_KERNEL_AUTO || want_queued == tuned->queued);
    if (use_tuned) {
        sh = *tuned;
    } else {
        sh.queued = kernel == ST_KERNEL_AUTO ? kDefaultQueued : want_queued;
        sh.block = block > 0 ? block : 256;
        sh.grid = grid > 0 ? grid : 0;  // 0: set below from the variant's residency
        sh.stack = bvh_depth <= 16 && sh.block >= 128 ? 16 : 32;
        sh.waves = sh.queued ? (sh.stack == 16 ? 5 : 3) : (sh.stack == 16 ? 6 : 4);
        // the queued form keeps the traversal stack in private memory where the
        // BVH fits it, which leaves LDS for 7 waves per SIMD: with candidate
        // lists BVH walks are rare; without them the walks go through the
        // scratch path, and the extra waves still win
        if (sh.queued && waves <= 0 && bvh_depth <= kStackDepth && trace_variant_exists(sh.block, 0, 7, true)) {
            sh.stack = 0;
            sh.waves = 7;
        }
    }
    if (waves > 0 && !use_tuned) {
        sh.waves = waves;
        for (int stk : {16, 32, 0})
            if (stack_entries(stk) >= bvh_depth && trace_variant_exists(sh.block, stk, waves, sh.queued)) {
                sh.stack = stk;
                break;
            }
    } else {
        if (stack_entries(sh.stack) < bvh_depth) sh.stack = 32;
        if (!trace_variant_exists(sh.block, sh.stack, sh.waves, sh.queued)) {
            sh.stack = 32;
            sh.waves = sh.queued ? 3 : 4;
            if (!trace_variant_exists(sh.block, sh.stack, sh.waves, sh.queued)) sh.block = 128;
        }
    }
    if (sh.queued && !want_grouped && min_elements <= kLdsHitElements && block <= 0 && waves <= 0 &&
        bvh_depth <= kStackDepth) {
        // the stack in private memory (7 waves per SIMD)
        sh.block = 256;
        sh.stack = 0;
        sh.waves = 7;
        sh.lds_hits = true;
        if (grid <= 0) sh.grid = 0;
    }
    // persistent grid: as many blocks as the device holds at once (a block
    // beyond that starts only when a resident one ends, near the launch's end,
    // and takes chunks there)
    if (sh.grid <= 0) {
        const int rw = trace_resident_waves(sh);
        sh.grid = rw > 0 ? rw / (sh.block / 64) : cus * 8;
    }
    // Chunks of 64 to 2048 rays. Every fetch from the shared chunk counter (a
    // single address, ~20 ns each in series) costs launch time, so a launch
    // that fits one chunk per resident wave gets exactly that (launch_trace
    // then skips the counter); larger ones get about four chunks per wave.
    const uint64_t resident = (uint64_t)trace_resident_waves(sh);
    const uint64_t grid_waves = (uint64_t)sh.grid * (uint64_t)(sh.block / 64);
    uint64_t chunk = resident > 0 && rays <= resident * 2048 ? (rays + resident - 1) / resident
                                                              : rays / (4 * grid_waves);
    chunk = (chunk + 63) / 64 * 64;
    sh.chunk = (int)std::min<uint64_t>(std::max<uint64_t>(chunk, 64), 2048);
    sh.kernel = sh.queued ? ST_KERNEL_QUEUED : ST_KERNEL_MEGA;
    return sh;
}

}  // namespace st

#define HIP_TRY(expr, what)                            \
    do {                                               \
        hipError_t e__ = (expr);                       \
        if (e__ != hipSuccess) return hip_fail(e__, what); \
    } while (0)

#define ST_TRY(expr)                 \
    do {                             \
        int rc__ = (expr);           \
        if (rc__ != ST_OK) return rc__; \
    } while (0)

// ---------------------------------------------------------------------------
extern "C" int st_abi_version(void) { return ST_ABI_VERSION; }
extern "C" const char* st_last_error(void) { return g_last_error.c_str(); }

extern "C" int st_device_count(int* count) {
    if (!count) return fail(ST_ERR_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        (void)hipGetLastError();
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return ST_OK;
}

extern "C" int st_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    int n = 0;
    ST_TRY(st_device_count(&n));
    if (device < 0 || device >= n) return fail(ST_ERR_INVALID, "device %d outside [0, %d)", device, n);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return ST_OK;
}

// ---------------------------------------------------------------------------
// Device scratch of st_trace_noalloc: a file-scope array, one copy per GPU for
// the life of the process. [0, kScratchCounterBytes) holds the trace's chunk
// counter, the rest the copies of a small flux map. Its device address is taken
// once per device (st_context_create, device current) and kept in
// g_scratch_ptr; host code passes that address, never the array's name.
namespace st {
constexpr size_t kScratchCounterBytes = 256;
constexpr size_t kScratchFluxBytes = 2u << 20;  // every map the flux-copy rule applies to fits
constexpr size_t kScratchBytes = kScratchCounterBytes + kScratchFluxBytes;
constexpr int kScratchDevices = 64;
}  // namespace st
static __device__ __align__(256) unsigned char scratch[st::kScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[st::kScratchDevices] = {};

namespace st {
static int scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return fail(ST_ERR_INVALID, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return ST_OK;
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(ST_ERR_DEVICE, "no device address for the trace scratch");
    g_scratch_ptr[device] = p;
    return ST_OK;
}
static void* scratch_of(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace st

extern "C" int st_context_create(int device, st_context* out) {
    if (!out) return fail(ST_ERR_INVALID, "null output handle");
    *out = nullptr;
    int n = 0;
    ST_TRY(st_device_count(&n));
    if (device < 0 || device >= n) return fail(ST_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    ST_TRY(scratch_setup(device));
    auto ctx = std::make_unique<st_context_s>();
    ctx->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (hipEventCreateWithFlags(&ctx->setup_done, hipEventDisableTiming) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        return fail(ST_ERR_DEVICE, "hipEventCreate failed");
    }
    *out = ctx.release();
    return ST_OK;
}

namespace st {
void context_retain(st_context ctx) { ctx->refs.fetch_add(1); }
void context_release(st_context ctx) {
    if (ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->setup_done);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int order_scene_traces(st_scene s, hipStream_t st) {
    if (!s->trace_done || s->trace_stream == st || capturing(st)) return ST_OK;
    HIP_TRY(hipStreamWaitEvent(st, s->trace_done, 0), "wait for the scene's last trace");
    return ST_OK;
}

int mark_scene_trace(st_scene s, hipStream_t st) {
    if (capturing(st)) return ST_OK;
    if (!s->trace_done) HIP_TRY(hipEventCreateWithFlags(&s->trace_done, hipEventDisableTiming), "trace event");
    HIP_TRY(hipEventRecord(s->trace_done, st), "trace event");
    s->trace_stream = st;
    return ST_OK;
}

st_launch_info launch_info(const LaunchShape& sh, uint64_t chunks, uint32_t copies) {
    st_launch_info li;
    memset(&li, 0, sizeof(li));
    li.kernel = !sh.queued ? ST_KERNEL_MEGA : sh.group_bvh ? ST_KERNEL_QUEUED_GROUPED : ST_KERNEL_QUEUED;
    li.block_size = sh.block;
    li.grid_blocks = sh.grid;
    li.stack = sh.stack;
    li.waves_per_simd = sh.waves;
    li.chunk_rays = sh.chunk;
    li.lds_hits = sh.lds_hits ? 1 : 0;
    li.flux_copies = (int32_t)copies;
    li.chunks = chunks;
    return li;
}

int order_before_trace(st_context ctx, hipStream_t user) {
    if (!user || user == ctx->stream || capturing(user)) return ST_OK;
    // nothing queued on the context stream: no dependency (a wait would cost
    // the caller's stream a barrier in front of the trace)
    if (hipStreamQuery(ctx->stream) == hipSuccess) return ST_OK;
    std::lock_guard<std::mutex> lk(ctx->order_mu);
    HIP_TRY(hipEventRecord(ctx->setup_done, ctx->stream), "set-up event");
    HIP_TRY(hipStreamWaitEvent(user, ctx->setup_done, 0), "wait for set-up");
    return ST_OK;
}

}  // namespace st

extern "C" int st_context_destroy(st_context ctx) {
    if (!ctx) return fail(ST_ERR_INVALID, "null context");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    context_release(ctx);
    return ST_OK;
}

extern "C" int st_buffer_alloc(st_context ctx, size_t bytes, void** dev_ptr) {
    if (!ctx || !dev_ptr) return fail(ST_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    return dev_alloc(dev_ptr, bytes, "user buffer");
}

extern "C" int st_buffer_free(st_context ctx, void* dev_ptr) {
    if (!ctx) return fail(ST_ERR_INVALID, "null context");
    if (!dev_ptr) return ST_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    HIP_TRY(hipFree(dev_ptr), "hipFree");
    return ST_OK;
}

extern "C" int st_buffer_upload(st_context ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ST_ERR_INVALID, "null argument");
    if (!bytes) return ST_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "upload");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ST_OK;
}

extern "C" int st_buffer_download(st_context ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ST_ERR_INVALID, "null argument");
    if (!bytes) return ST_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipDeviceSynchronize(), "sync");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ST_OK;
}

extern "C" int st_buffer_zero(st_context ctx, void* dev_ptr, size_t bytes) {
    if (!ctx || (!dev_ptr && bytes)) return fail(ST_ERR_INVALID, "null argument");
    if (!bytes) return ST_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipMemsetAsync(dev_ptr, 0, bytes, ctx->stream), "memset");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ST_OK;
}

// ---------------------------------------------------------------------------
static int validate_elements(const st_element_f64* el, int32_t n) {
    if (!el) return fail(ST_ERR_INVALID, "null element array");
    if (n <= 0) return fail(ST_ERR_INVALID, "scene needs at least one element (got %d)", n);
    if ((uint32_t)n >= kMaxElements) return fail(ST_ERR_INVALID, "too many elements (%d)", n);
    for (int32_t i = 0; i < n; ++i) {
        const st_element_f64& e = el[i];
        if (e.aperture != ST_APERTURE_RECTANGLE)
            return fail(ST_ERR_INVALID, "element %d: only rectangular apertures are supported", i);
        if (e.surface != ST_SURFACE_FLAT && e.surface != ST_SURFACE_PARABOLIC && e.surface != ST_SURFACE_CYLINDER)
            return fail(ST_ERR_INVALID, "element %d: unknown surface type %d", i, e.surface);
        if (e.role != ST_ROLE_MIRROR && e.role != ST_ROLE_RECEIVER)
            return fail(ST_ERR_INVALID, "element %d: unknown role %d", i, e.role);
        if (!(e.dim_x > 0) || !(e.dim_y > 0)) return fail(ST_ERR_INVALID, "element %d: non-positive aperture", i);
        if (!(e.reflectivity >= 0 && e.reflectivity <= 1))
            return fail(ST_ERR_INVALID, "element %d: reflectivity outside [0,1]", i);
        if (!(e.slope_error >= 0) || !(e.spec_error >= 0))
            return fail(ST_ERR_INVALID, "element %d: negative optical error", i);
        double d[3] = {e.aim_point[0] - e.origin[0], e.aim_point[1] - e.origin[1], e.aim_point[2] - e.origin[2]};
        if (!(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > 0))
            return fail(ST_ERR_INVALID, "element %d: aim point equals origin", i);
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(e.origin[k]) || !std::isfinite(e.aim_point[k]))
                return fail(ST_ERR_INVALID, "element %d: non-finite coordinates", i);
    }
    return ST_OK;
}

namespace st {
// frames + refit for the scene's current elements / aim overrides (async on ctx stream)
int scene_refresh(st_scene s, bool use_overrides) {
    hipStream_t st = s->ctx->stream;
    if (launch_injected()) return fail(ST_ERR_INJECTED, "injected launch failure (frames)");
    HIP_TRY(launch_frames(st, (const st_element_f64*)s->d_elems.ptr, s->n, (const int32_t*)s->d_elem_slot.ptr,
                          use_overrides ? (const double*)s->d_aim.ptr : nullptr,
                          use_overrides ? (const double*)s->d_zrot.ptr : nullptr, (double*)s->d_frames.ptr,
                          (double*)s->d_ref_aabb.ptr, (float*)s->d_bounds.ptr, (float4*)s->d_prims.ptr,
                          (uint32_t*)s->d_meta.ptr, (float4*)s->d_optics.ptr),
            "frames kernel");
    HIP_TRY(launch_refit(st, (const int32_t*)s->d_child.ptr, (const int32_t*)s->d_level_nodes.ptr,
                         s->level_offsets.data(), (int32_t)s->level_offsets.size() - 1, s->n_nodes, s->n, 1,
                         (const float*)s->d_bounds.ptr, (float4*)s->d_nodes.ptr),
            "refit kernel");
    return ST_OK;
}

int scene_sun_plane(st_scene s) {
    if (!s->sun_set) return ST_OK;
    hipStream_t st = s->ctx->stream;
    HIP_TRY(hipMemcpyAsync(s->d_sun_unit.ptr, s->sun_unit, sizeof(double) * 3, hipMemcpyHostToDevice, st),
            "sun vector upload");
    HIP_TRY(launch_sun_plane(st, (const double*)s->d_ref_aabb.ptr, s->n, 1, (const double*)s->d_sun_unit.ptr,
                             &s->sun, (SunFrame*)s->d_sun.ptr),
            "sun plane kernel");
    SunFrame F;
    HIP_TRY(hipMemcpyAsync(&F, s->d_sun.ptr, sizeof(F), hipMemcpyDeviceToHost, st), "sun plane download");
    HIP_TRY(hipStreamSynchronize(st), "sun plane sync");
    for (int k = 0; k < 3; ++k) {
        s->plane.v0[k] = F.v0[k];
        s->plane.v1[k] = F.v1[k];
        s->plane.v2[k] = F.v2[k];
        s->plane.v3[k] = F.v3[k];
    }
    s->plane.distance = F.distance;
    s->plane.area = F.area;
    s->plane.box_half_angle = F.box_half_angle;
    // reflection lists cost more to build than the grid: rebuilt on demand
    s->bounce_valid = s->pair_valid = false;
    s->lists_stale = true;
    s->rays_since_change = 0;
    return scene_grid(s);
}

// Reflection candidate lists are (re)built once enough rays are traced on the
// current scene state to repay the build; until then reflected rays traverse
// the BVH. The threshold follows the scene's last measured build time at
// kListSavingPerRay seconds saved per traced ray.
int scene_ensure_lists(st_scene s, uint64_t rays, bool force) {
    if (!s->lists_stale || !s->sun_set) return ST_OK;
    s->rays_since_change += rays;
    if (!force && s->rays_since_change < s->list_min_rays) return ST_OK;
    const auto t0 = std::chrono::steady_clock::now();
    ST_TRY(scene_bounce_lists(s));
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    s->list_min_rays = (uint64_t)std::min(4.0e9, std::max((double)kListMinRays, dt / kListSavingPerRay));
    s->lists_stale = false;
    return ST_OK;
}

static double sun_cone_tan(const st_sun_desc& sun) { return tan(sun_cone_angle(sun)); }

// walk pass (lists into per-task scratch) -> host prefix sum (lists over
// kMaxBounceList or marked none get an empty-but-invalid range) -> compaction.
template <class Launch>
static int build_cone_lists(st_scene s, size_t tasks, DevBuf& counts, DevBuf& range, DevBuf& items,
                            std::vector<uint2>& h_range, uint64_t* total_out, Launch launch) {
    hipStream_t st = s->ctx->stream;
    *total_out = 0;
    if (tasks == 0) return ST_OK;
    ST_TRY(counts.ensure(sizeof(uint32_t) * tasks, "list counts"));
    ST_TRY(range.ensure(sizeof(uint2) * tasks, "list ranges"));
    ST_TRY(s->d_list_scratch.ensure(sizeof(int32_t) * kMaxBounceList * tasks, "list scratch"));
    HIP_TRY(launch((uint32_t*)counts.ptr, (int32_t*)s->d_list_scratch.ptr), "list walk kernel");
    s->h_counts.resize(tasks);
    h_range.resize(tasks);
    HIP_TRY(hipMemcpyAsync(s->h_counts.data(), counts.ptr, sizeof(uint32_t) * tasks, hipMemcpyDeviceToHost, st),
            "list counts download");
    HIP_TRY(hipStreamSynchronize(st), "list sync");
    uint64_t total = 0;
    for (size_t t = 0; t < tasks; ++t) {
        const uint32_t c = s->h_counts[t];
        if (c == kNoBounceList || total + c > 0x7FFFFFFFull) {
            h_range[t] = make_uint2(1u, 0u);
        } else {
            h_range[t] = make_uint2((uint32_t)total, (uint32_t)(total + c));
            total += c;
        }
    }
    *total_out = total;
    if (total == 0) return ST_OK;
    ST_TRY(items.ensure(sizeof(int32_t) * total, "list items"));
    HIP_TRY(hipMemcpyAsync(range.ptr, h_range.data(), sizeof(uint2) * tasks, hipMemcpyHostToDevice, st),
            "list ranges upload");
    HIP_TRY(launch_list_compact(st, (int32_t)tasks, (const uint2*)range.ptr, (const int32_t*)s->d_list_scratch.ptr,
                                (int32_t*)items.ptr),
            "list compaction kernel");
    HIP_TRY(hipStreamSynchronize(st), "list sync");  // h_range and the scratch are reused by the next pass
    return ST_OK;
}

int scene_bounce_lists(st_scene s) {
    hipStream_t st = s->ctx->stream;
    ++s->list_builds;
    s->bounce_valid = s->pair_valid = false;
    const double theta = sun_cone_angle(s->sun);
    // mirrors aimed by the bisector rule at a target for the sun traced: the
    // first-reflection lists test the re-aim candidates
    const bool reaim = s->aim_bisector && s->aim_sun[0] == s->sun_unit[0] && s->aim_sun[1] == s->sun_unit[1] &&
                       s->aim_sun[2] == s->sun_unit[2];
    if (reaim) ST_TRY(ensure_reaim(s->reaim, s, s->aim_target, theta, st));
    auto common = [&](int32_t tasks, const int32_t* pitems, const int32_t* powner, uint32_t* c, int32_t* scratch) {
        const bool sup = reaim && !pitems;
        return launch_cone_lists(st, (const st_element_f64*)s->d_elems.ptr, tasks, tasks, s->n, s->n_nodes,
                                 (const int32_t*)s->d_elem_slot.ptr, (const uint32_t*)s->d_meta.ptr,
                                 (const double*)s->d_frames.ptr, (const float*)s->d_bounds.ptr,
                                 (const float4*)s->d_nodes.ptr, (const SunFrame*)s->d_sun.ptr, theta, pitems, powner,
                                 c, nullptr, scratch, sup ? (const uint2*)s->reaim.range.ptr : nullptr,
                                 sup ? (const int32_t*)s->reaim.items.ptr : nullptr);
    };
    // first reflections: one list per element
    ST_TRY(build_cone_lists(s, (size_t)s->n, s->d_bounce_counts, s->d_bounce_range, s->d_bounce_items,
                            s->h_bounce_range, &s->bounce_items,
                            [&](uint32_t* c, int32_t* scratch) { return common(s->n, nullptr, nullptr, c, scratch); }));
    if (s->bounce_items == 0) return ST_OK;
    s->bounce_valid = true;
    // second reflections: one list per first-reflection list entry
    s->h_pair_owner.assign(s->bounce_items, 0);
    for (int32_t e = 0; e < s->n; ++e)  // a list's first entry is its header: no second-reflection task
        for (uint32_t j = s->h_bounce_range[e].x; j < s->h_bounce_range[e].y; ++j)
            s->h_pair_owner[j] = j == s->h_bounce_range[e].x ? -1 : e;
    ST_TRY(s->d_pair_owner.ensure(sizeof(int32_t) * s->bounce_items, "pair owners"));
    HIP_TRY(hipMemcpyAsync(s->d_pair_owner.ptr, s->h_pair_owner.data(), sizeof(int32_t) * s->bounce_items,
                           hipMemcpyHostToDevice, st),
            "pair owners upload");
    const int32_t tasks = (int32_t)std::min<uint64_t>(s->bounce_items, 0x7FFFFFFF);
    ST_TRY(build_cone_lists(s, (size_t)tasks, s->d_pair_counts, s->d_pair_range, s->d_pair_items, s->h_pair_range,
                            &s->pair_items, [&](uint32_t* c, int32_t* scratch) {
                                return common(tasks, (const int32_t*)s->d_bounce_items.ptr,
                                              (const int32_t*)s->d_pair_owner.ptr, c, scratch);
                            }));
    s->pair_valid = s->pair_items > 0;
    // share of the mirrors in first-reflection lists whose pair list is missing
    // (too long or too wide: those second reflections walk the BVH)
    std::vector<int32_t> items((size_t)tasks);
    HIP_TRY(hipMemcpyAsync(items.data(), s->d_bounce_items.ptr, sizeof(int32_t) * items.size(), hipMemcpyDeviceToHost,
                           st),
            "list items download");
    HIP_TRY(hipStreamSynchronize(st), "list sync");
    uint64_t entries = 0, missing = 0;
    for (int32_t t = 0; t < tasks; ++t) {
        if (s->h_pair_owner[t] < 0) continue;
        const int32_t e = s->topo.slot_elem[items[t]];
        if (s->host_elems[e].role != ST_ROLE_MIRROR) continue;
        ++entries;
        missing += s->h_pair_range[t].x > s->h_pair_range[t].y;
    }
    s->pair_missing = entries ? (double)missing / (double)entries : 1.0;
    if (getenv("ST_LIST_LOG"))
        fprintf(stderr, "lists: first %llu items, pairs %llu items, %llu of %llu entries without a pair list\n",
                (unsigned long long)s->bounce_items, (unsigned long long)s->pair_items, (unsigned long long)missing,
                (unsigned long long)entries);
    return ST_OK;
}

// Whether a trace's reflected rays mix list segments with BVH ones, for the
// grouped form: whenever first-reflection lists exist (rays after a second
// mirror always walk the BVH, as do those whose pair list is missing). 10k
// field, 1e8 rays, block 256 with the stack in private memory, one box: the
// grouped form takes -8 / -19 / -22 / -23 / -11 / -1 / -1 / -1.5 % of the plain
// queued form's time at 5 / 15 / 25 / 30 / 40 / 50 / 60 / 80 deg elevation.
static bool mixed_bvh_segments(st_scene, const JobDev& J) { return J.bounce_range != nullptr; }

int ensure_reaim(ReaimCache& c, st_scene s, const double target[3], double theta, hipStream_t st) {
    if (c.valid && c.epoch == s->geom_epoch && c.theta == theta && c.target[0] == target[0] &&
        c.target[1] == target[1] && c.target[2] == target[2])
        return ST_OK;
    c.valid = false;
    const size_t n = (size_t)s->n;
    ST_TRY(c.spheres.ensure(sizeof(double4) * n, "re-aim spheres"));
    ST_TRY(c.range.ensure(sizeof(uint2) * n, "re-aim list ranges"));
    ST_TRY(c.items.ensure(sizeof(int32_t) * kMaxBounceList * n, "re-aim lists"));
    ST_TRY(c.target_dev.ensure(sizeof(double) * 3, "re-aim target"));
    for (int k = 0; k < 3; ++k) c.target[k] = target[k];
    HIP_TRY(hipMemcpyAsync(c.target_dev.ptr, c.target, sizeof(double) * 3, hipMemcpyHostToDevice, st), "upload");
    HIP_TRY(launch_reaim_lists(st, (const st_element_f64*)s->d_elems.ptr, s->n, (const int32_t*)s->d_elem_slot.ptr,
                               (const uint32_t*)s->d_meta.ptr, (const double*)c.target_dev.ptr, theta,
                               (double4*)c.spheres.ptr, (uint2*)c.range.ptr, (int32_t*)c.items.ptr),
            "re-aim candidate lists");
    c.theta = theta;
    c.epoch = s->geom_epoch;
    c.valid = true;
    return ST_OK;
}

int route_flux_copies(DevBuf& buf, JobDev* jobs, int32_t n, uint64_t bins, uint64_t rays, hipStream_t st,
                      uint32_t* copies) {
    uint32_t c = flux_copies(bins, rays);
    while (c > 1 && (uint64_t)n * c * bins * sizeof(int64_t) > kFluxCopyBytes) c >>= 1;
    *copies = c;
    if (c == 1) return ST_OK;
    const size_t need = (size_t)n * c * bins * sizeof(int64_t);
    if (buf.bytes < need) {
        // fresh memory: cleared once; every gather leaves the copies zero
        ST_TRY(buf.ensure(need, "flux copies"));
        HIP_TRY(launch_fill_u64(st, buf.ptr, buf.bytes / sizeof(int64_t), 0), "clear flux copies");
    }
    for (int32_t k = 0; k < n; ++k) {
        if (!jobs[k].flux) continue;
        jobs[k].flux = (int64_t*)buf.ptr + (size_t)k * c * bins;
        jobs[k].flux_copy_mask = c - 1;
    }
    return ST_OK;
}

int scene_grid(st_scene s) {
    hipStream_t st = s->ctx->stream;
    s->grid_valid = false;
    if (s->n < kGridMinElements) return ST_OK;
    ST_TRY(s->d_grid_counts.ensure(sizeof(uint32_t) * kGridCells, "grid counts"));
    ST_TRY(s->d_grid_cursor.ensure(sizeof(uint32_t) * kGridCells, "grid cursor"));
    ST_TRY(s->d_grid_start.ensure(sizeof(uint32_t) * (kGridCells + 1), "grid start"));
    ST_TRY(s->d_grid_stats.ensure(sizeof(uint32_t) * 2, "grid stats"));
    const double tmax = sun_cone_tan(s->sun);
    HIP_TRY(launch_fill_u64(st, s->d_grid_counts.ptr, kGridCells / 2, 0), "grid clear");
    HIP_TRY(launch_grid_count(st, (const float*)s->d_bounds.ptr, (const float4*)s->d_prims.ptr,
                              (const uint32_t*)s->d_meta.ptr, s->n, 1, (const SunFrame*)s->d_sun.ptr, tmax,
                              (uint32_t*)s->d_grid_counts.ptr, (uint32_t*)s->d_grid_start.ptr,
                              (uint32_t*)s->d_grid_stats.ptr, nullptr, nullptr, 0),
            "grid count kernels");
    uint32_t stats[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(stats, s->d_grid_stats.ptr, sizeof(stats), hipMemcpyDeviceToHost, st), "grid stats");
    HIP_TRY(hipStreamSynchronize(st), "grid sync");
    s->grid_items = stats[0];
    s->grid_longest = stats[1];
    if (!grid_lists_pay(stats[0], stats[1])) return ST_OK;
    ST_TRY(s->d_grid_items.ensure(sizeof(int32_t) * stats[0], "grid items"));
    HIP_TRY(launch_fill_u64(st, s->d_grid_cursor.ptr, kGridCells / 2, 0), "grid cursor clear");
    HIP_TRY(launch_grid_fill(st, (const float*)s->d_bounds.ptr, (const float4*)s->d_prims.ptr,
                             (const uint32_t*)s->d_meta.ptr, s->n, 1, (const SunFrame*)s->d_sun.ptr, tmax,
                             (const uint32_t*)s->d_grid_start.ptr, (uint32_t*)s->d_grid_cursor.ptr,
                             (int32_t*)s->d_grid_items.ptr, stats[0]),
            "grid fill kernel");
    s->grid_valid = true;
    return ST_OK;
}
}  // namespace st

extern "C" int st_scene_create(st_context ctx, const st_element_f64* elements, int32_t n, st_scene* out) {
    if (!ctx || !out) return fail(ST_ERR_INVALID, "null context or output handle");
    *out = nullptr;
    ST_TRY(validate_elements(elements, n));
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    auto s = std::make_unique<st_scene_s>();
    s->ctx = ctx;
    s->ctx_ref.hold(ctx);
    s->n = n;
    s->host_elems.assign(elements, elements + n);
    // host FP64 bounds -> BVH topology
    std::vector<float> lo_hi(6 * (size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        double l2g[9];
        element_frame(elements[i], elements[i].aim_point, elements[i].zrot_deg, l2g);
        element_bounds(elements[i], l2g, &lo_hi[6 * i], &lo_hi[6 * i + 3]);
    }
    try {
        s->topo = build_bvh(lo_hi.data(), n, kMaxLeaf);
    } catch (const std::exception& ex) {
        return fail(ST_ERR_INVALID, "BVH build failed: %s", ex.what());
    }
    s->n_nodes = s->topo.n_nodes();
    // nodes grouped by level
    const int32_t levels = s->topo.depth;
    s->level_offsets.assign(levels + 1, 0);
    for (int32_t lv : s->topo.node_level) s->level_offsets[lv + 1]++;
    for (int32_t l = 0; l < levels; ++l) s->level_offsets[l + 1] += s->level_offsets[l];
    std::vector<int32_t> level_nodes(s->n_nodes), fillp(s->level_offsets.begin(), s->level_offsets.end() - 1);
    for (int32_t i = 0; i < s->n_nodes; ++i) level_nodes[fillp[s->topo.node_level[i]]++] = i;
    // device buffers
    const size_t N = (size_t)n, NN = (size_t)s->n_nodes;
    ST_TRY(s->d_elems.ensure(sizeof(st_element_f64) * N, "elements"));
    ST_TRY(s->d_elem_slot.ensure(sizeof(int32_t) * N, "element slots"));
    ST_TRY(s->d_frames.ensure(sizeof(double) * 9 * N, "frames"));
    ST_TRY(s->d_ref_aabb.ensure(sizeof(double) * 6 * N, "reference AABBs"));
    ST_TRY(s->d_bounds.ensure(sizeof(float) * 6 * N, "bounds"));
    ST_TRY(s->d_prims.ensure(sizeof(float4) * 4 * N, "primitives"));
    ST_TRY(s->d_meta.ensure(sizeof(uint32_t) * N, "primitive meta"));
    ST_TRY(s->d_optics.ensure(sizeof(float4) * N, "optics"));
    ST_TRY(s->d_nodes.ensure(sizeof(float4) * 4 * NN, "BVH nodes"));
    ST_TRY(s->d_child.ensure(sizeof(int32_t) * 2 * NN, "BVH children"));
    ST_TRY(s->d_level_nodes.ensure(sizeof(int32_t) * NN, "BVH levels"));
    ST_TRY(s->d_sun.ensure(sizeof(SunFrame), "sun frame"));
    ST_TRY(s->d_sun_unit.ensure(sizeof(double) * 3, "sun vector"));
    ST_TRY(s->d_aim.ensure(sizeof(double) * std::max<size_t>(3 * N, 6), "aim points"));
    ST_TRY(s->d_zrot.ensure(sizeof(double) * N, "zrot"));
    ST_TRY(s->d_counter.ensure(sizeof(uint32_t) * 2, "work counter"));
    hipStream_t st = ctx->stream;
    HIP_TRY(hipMemcpyAsync(s->d_elems.ptr, elements, sizeof(st_element_f64) * N, hipMemcpyHostToDevice, st), "upload");
    HIP_TRY(hipMemcpyAsync(s->d_elem_slot.ptr, s->topo.elem_slot.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st),
            "upload");
    HIP_TRY(hipMemcpyAsync(s->d_child.ptr, s->topo.child.data(), sizeof(int32_t) * 2 * NN, hipMemcpyHostToDevice, st),
            "upload");
    HIP_TRY(hipMemcpyAsync(s->d_level_nodes.ptr, level_nodes.data(), sizeof(int32_t) * NN, hipMemcpyHostToDevice, st),
            "upload");
    HIP_TRY(hipMemsetAsync(s->d_counter.ptr, 0, sizeof(uint32_t) * 2, st), "counter clear");
    ST_TRY(scene_refresh(s.get(), false));
    HIP_TRY(hipStreamSynchronize(st), "scene build sync");
    *out = s.release();
    return ST_OK;
}

st_scene_s::~st_scene_s() {
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (trace_done) {
        (void)hipEventSynchronize(trace_done);
        (void)hipEventDestroy(trace_done);
    }
    if (batch_graph_exec) (void)hipGraphExecDestroy(batch_graph_exec);
}

extern "C" int st_scene_destroy(st_scene scene) {
    if (!scene) return fail(ST_ERR_INVALID, "null scene");
    delete scene;  // device buffers are freed before the context reference goes
    return ST_OK;
}

extern "C" int st_scene_num_elements(st_scene scene, int32_t* n) {
    if (!scene || !n) return fail(ST_ERR_INVALID, "null argument");
    *n = scene->n;
    return ST_OK;
}

extern "C" int st_scene_update_aim(st_scene s, const double* aim_points, const double* zrot_deg) {
    if (!s || !aim_points) return fail(ST_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    for (int32_t i = 0; i < s->n; ++i) {
        double d[3] = {aim_points[3 * i] - s->host_elems[i].origin[0], aim_points[3 * i + 1] - s->host_elems[i].origin[1],
                       aim_points[3 * i + 2] - s->host_elems[i].origin[2]};
        if (!(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > 0)) return fail(ST_ERR_INVALID, "element %d: aim point equals origin", i);
        for (int k = 0; k < 3; ++k) s->host_elems[i].aim_point[k] = aim_points[3 * i + k];
        if (zrot_deg) s->host_elems[i].zrot_deg = zrot_deg[i];
    }
    s->aim_bisector = false;
    hipStream_t st = s->ctx->stream;
    HIP_TRY(hipMemcpyAsync(s->d_elems.ptr, s->host_elems.data(), sizeof(st_element_f64) * s->n, hipMemcpyHostToDevice, st),
            "upload");
    ST_TRY(scene_refresh(s, false));
    ST_TRY(scene_sun_plane(s));
    HIP_TRY(hipStreamSynchronize(st), "update sync");
    return ST_OK;
}

extern "C" int st_scene_update_elements(st_scene s, const st_element_f64* elements, int32_t n) {
    if (!s) return fail(ST_ERR_INVALID, "null scene");
    if (n != s->n) return fail(ST_ERR_INVALID, "element count changed (%d -> %d); create a new scene", s->n, n);
    ST_TRY(validate_elements(elements, n));
    for (int32_t i = 0; i < n; ++i)
        if (elements[i].surface != s->host_elems[i].surface || elements[i].role != s->host_elems[i].role)
            return fail(ST_ERR_INVALID, "element %d changed type or role; create a new scene", i);
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    s->host_elems.assign(elements, elements + n);
    s->aim_bisector = false;
    ++s->geom_epoch;
    hipStream_t st = s->ctx->stream;
    HIP_TRY(hipMemcpyAsync(s->d_elems.ptr, elements, sizeof(st_element_f64) * n, hipMemcpyHostToDevice, st), "upload");
    ST_TRY(scene_refresh(s, false));
    ST_TRY(scene_sun_plane(s));
    HIP_TRY(hipStreamSynchronize(st), "update sync");
    return ST_OK;
}

namespace st {
static int check_sun(const st_sun_desc* sun, double* len) {
    const double l = std::sqrt(sun->vector[0] * sun->vector[0] + sun->vector[1] * sun->vector[1] +
                               sun->vector[2] * sun->vector[2]);
    if (!(l > 0) || !std::isfinite(l)) return fail(ST_ERR_INVALID, "sun vector must be finite and non-zero");
    if (sun->shape != ST_SUN_PILLBOX && sun->shape != ST_SUN_GAUSSIAN) return fail(ST_ERR_INVALID, "unknown sun shape %d", sun->shape);
    if (!(sun->param >= 0) || sun->param > 0.2) return fail(ST_ERR_INVALID, "sun angle %g rad outside [0, 0.2]", sun->param);
    if (sun->position_sampling != ST_POS_PHILOX && sun->position_sampling != ST_POS_HALTON &&
        sun->position_sampling != ST_POS_STRATIFIED)
        return fail(ST_ERR_INVALID, "unknown position sampling %d", sun->position_sampling);
    *len = l;
    return ST_OK;
}
static void store_sun(st_scene s, const st_sun_desc* sun, double l) {
    s->sun = *sun;
    if (!(s->sun.dni > 0)) s->sun.dni = 1000.0;
    for (int k = 0; k < 3; ++k) s->sun_unit[k] = sun->vector[k] / l;
    s->sun_set = true;
}
}  // namespace st

extern "C" int st_scene_update(st_scene s, const st_element_f64* elements, int32_t n, const st_sun_desc* sun) {
    if (!s) return fail(ST_ERR_INVALID, "null scene");
    if (!elements && !sun) return fail(ST_ERR_INVALID, "nothing to update");
    if (elements) {
        if (n != s->n) return fail(ST_ERR_INVALID, "element count changed (%d -> %d); create a new scene", s->n, n);
        ST_TRY(validate_elements(elements, n));
        for (int32_t i = 0; i < n; ++i)
            if (elements[i].surface != s->host_elems[i].surface || elements[i].role != s->host_elems[i].role)
                return fail(ST_ERR_INVALID, "element %d changed type or role; create a new scene", i);
    }
    double l = 0;
    if (sun) ST_TRY(check_sun(sun, &l));
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    hipStream_t st = s->ctx->stream;
    if (elements) {
        s->host_elems.assign(elements, elements + n);
        s->aim_bisector = false;
        ++s->geom_epoch;
        HIP_TRY(hipMemcpyAsync(s->d_elems.ptr, elements, sizeof(st_element_f64) * n, hipMemcpyHostToDevice, st),
                "upload");
        ST_TRY(scene_refresh(s, false));
    }
    if (sun) store_sun(s, sun, l);
    ST_TRY(scene_sun_plane(s));  // once, for the new poses and sun
    HIP_TRY(hipStreamSynchronize(st), "update sync");
    return ST_OK;
}

extern "C" int st_scene_aim_heliostats(st_scene s, const double sun_vector[3], const double target[3]) {
    if (!s || !sun_vector || !target) return fail(ST_ERR_INVALID, "null argument");
    double l = std::sqrt(sun_vector[0] * sun_vector[0] + sun_vector[1] * sun_vector[1] + sun_vector[2] * sun_vector[2]);
    if (!(l > 0)) return fail(ST_ERR_INVALID, "zero sun vector");
    double st_in[6] = {sun_vector[0] / l, sun_vector[1] / l, sun_vector[2] / l, target[0], target[1], target[2]};
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    hipStream_t st = s->ctx->stream;
    // sun unit vector and target share the scene's aim scratch (6 doubles)
    HIP_TRY(hipMemcpyAsync(s->d_aim.ptr, st_in, sizeof(st_in), hipMemcpyHostToDevice, st), "upload");
    HIP_TRY(launch_aim(st, (st_element_f64*)s->d_elems.ptr, s->n, (const double*)s->d_aim.ptr,
                       (const double*)s->d_aim.ptr + 3),
            "aim kernel");
    s->aim_bisector = true;
    for (int k = 0; k < 3; ++k) {
        s->aim_target[k] = target[k];
        s->aim_sun[k] = st_in[k];
    }
    ST_TRY(scene_refresh(s, false));
    ST_TRY(scene_sun_plane(s));  // synchronizes the stream
    HIP_TRY(hipMemcpyAsync(s->host_elems.data(), s->d_elems.ptr, sizeof(st_element_f64) * s->n, hipMemcpyDeviceToHost, st),
            "download");
    HIP_TRY(hipStreamSynchronize(st), "aim sync");
    return ST_OK;
}


extern "C" int st_scene_set_sun_and_aim(st_scene s, const st_sun_desc* sun, const double target[3]) {
    if (!s || !sun || !target) return fail(ST_ERR_INVALID, "null argument");
    double l = 0;
    ST_TRY(check_sun(sun, &l));
    store_sun(s, sun, l);
    const double st_in[6] = {s->sun_unit[0], s->sun_unit[1], s->sun_unit[2], target[0], target[1], target[2]};
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    hipStream_t st = s->ctx->stream;
    HIP_TRY(hipMemcpyAsync(s->d_aim.ptr, st_in, sizeof(st_in), hipMemcpyHostToDevice, st), "upload");
    HIP_TRY(launch_aim(st, (st_element_f64*)s->d_elems.ptr, s->n, (const double*)s->d_aim.ptr,
                       (const double*)s->d_aim.ptr + 3),
            "aim kernel");
    s->aim_bisector = true;
    for (int k = 0; k < 3; ++k) {
        s->aim_target[k] = target[k];
        s->aim_sun[k] = s->sun_unit[k];
    }
    ST_TRY(scene_refresh(s, false));
    ST_TRY(scene_sun_plane(s));  // one sun plane and grid, for the new sun and aims
    HIP_TRY(hipMemcpyAsync(s->host_elems.data(), s->d_elems.ptr, sizeof(st_element_f64) * s->n, hipMemcpyDeviceToHost, st),
            "download");
    HIP_TRY(hipStreamSynchronize(st), "aim sync");
    return ST_OK;
}

extern "C" int st_scene_set_sun(st_scene s, const st_sun_desc* sun) {
    if (!s || !sun) return fail(ST_ERR_INVALID, "null argument");
    double l = 0;
    ST_TRY(check_sun(sun, &l));
    store_sun(s, sun, l);
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    return scene_sun_plane(s);
}

extern "C" int st_scene_get_sun_plane(st_scene s, st_sun_plane* out) {
    if (!s || !out) return fail(ST_ERR_INVALID, "null argument");
    if (!s->sun_set) return fail(ST_ERR_STATE, "sun not set");
    *out = s->plane;
    return ST_OK;
}

extern "C" int st_scene_get_frames(st_scene s, double* l2g) {
    if (!s || !l2g) return fail(ST_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(l2g, s->d_frames.ptr, sizeof(double) * 9 * s->n, hipMemcpyDeviceToHost, s->ctx->stream), "download");
    HIP_TRY(hipStreamSynchronize(s->ctx->stream), "sync");
    return ST_OK;
}

extern "C" int st_scene_get_aabbs(st_scene s, float* lo_hi) {
    if (!s || !lo_hi) return fail(ST_ERR_INVALID, "null argument");
    std::vector<double> tmp(6 * (size_t)s->n);
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(tmp.data(), s->d_ref_aabb.ptr, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, s->ctx->stream),
            "download");
    HIP_TRY(hipStreamSynchronize(s->ctx->stream), "sync");
    for (size_t i = 0; i < tmp.size(); ++i) lo_hi[i] = (float)tmp[i];
    return ST_OK;
}

extern "C" int st_scene_bvh_info(st_scene s, int32_t* n_nodes, int32_t* depth, int32_t* max_leaf) {
    if (!s) return fail(ST_ERR_INVALID, "null scene");
    if (n_nodes) *n_nodes = s->n_nodes;
    if (depth) *depth = s->topo.depth;
    if (max_leaf) *max_leaf = s->topo.max_leaf;
    return ST_OK;
}

extern "C" int st_scene_synchronize(st_scene s) {
    if (!s) return fail(ST_ERR_INVALID, "null scene");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    HIP_TRY(hipStreamSynchronize(s->ctx->stream), "sync");
    return ST_OK;
}

// ---------------------------------------------------------------------------
namespace st {
int resolve_flux_element(st_scene s, int32_t requested, int32_t nx, int32_t ny, int32_t* slot) {
    *slot = -1;
    if (nx < 0 || ny < 0 || (int64_t)nx * ny > (1 << 26)) return fail(ST_ERR_INVALID, "bad flux grid %dx%d", nx, ny);
    if (nx == 0 || ny == 0) return ST_OK;
    int32_t fe = requested;
    if (fe < 0)
        for (int32_t i = s->n - 1; i >= 0; --i)
            if (s->host_elems[i].role == ST_ROLE_RECEIVER) { fe = i; break; }
    if (fe < 0) return ST_OK;  // no receiver: nothing to bin
    if (fe >= s->n) return fail(ST_ERR_INVALID, "flux element %d out of range", fe);
    if (s->host_elems[fe].role != ST_ROLE_RECEIVER) return fail(ST_ERR_INVALID, "flux element %d is not a receiver", fe);
    *slot = s->topo.elem_slot[fe];
    return ST_OK;
}

int fill_job(st_scene s, const st_trace_desc* d, const st_outputs* out, const void* prims, const void* nodes,
             const void* sun, JobDev* J) {
    memset(J, 0, sizeof(*J));
    J->prims = (const float4*)prims;
    J->meta = (const uint32_t*)s->d_meta.ptr;
    J->optics = (const float4*)s->d_optics.ptr;
    J->nodes = (const float4*)nodes;
    J->sun = (const SunFrame*)sun;
    if (s->grid_valid && prims == s->d_prims.ptr && sun == s->d_sun.ptr && !(d->flags & ST_TRACE_BVH_ONLY)) {
        J->grid_start = (const uint32_t*)s->d_grid_start.ptr;
        J->grid_items = (const int32_t*)s->d_grid_items.ptr;
    }
    if (s->bounce_valid && prims == s->d_prims.ptr && sun == s->d_sun.ptr && !(d->flags & ST_TRACE_BVH_ONLY)) {
        J->bounce_range = (const uint2*)s->d_bounce_range.ptr;
        J->bounce_items = (const int32_t*)s->d_bounce_items.ptr;
        if (s->pair_valid) {
            J->pair_range = (const uint2*)s->d_pair_range.ptr;
            J->pair_items = (const int32_t*)s->d_pair_items.ptr;
        }
    }
    J->ray_begin = d->ray_begin;
    J->ray_count = d->ray_count ? d->ray_count : d->num_rays - d->ray_begin;
    J->record_base = d->ray_begin;
    J->seed_lo = (uint32_t)d->seed;
    J->seed_hi = (uint32_t)(d->seed >> 32);
    J->job_id = d->job_id;
    J->max_depth = d->max_depth;
    J->flags = d->flags;
    J->flux_nx = d->flux_nx;
    J->flux_ny = d->flux_ny;
    J->flux_bins = (uint32_t)std::max<int64_t>((int64_t)d->flux_nx * d->flux_ny, 0);
    J->flux_copy_mask = 0;
    ST_TRY(resolve_flux_element(s, d->flux_element, d->flux_nx, d->flux_ny, &J->flux_slot));
    // the hottest counter is the main receiver's: aggregate it in registers
    // whether or not a flux map is requested (one atomic per wave and chunk)
    J->agg_slot = J->flux_slot;
    if (J->agg_slot < 0)
        for (int32_t i = s->n - 1; i >= 0; --i)
            if (s->host_elems[i].role == ST_ROLE_RECEIVER) {
                J->agg_slot = s->topo.elem_slot[i];
                break;
            }
    if (out) {
        J->flux = J->flux_slot >= 0 ? out->flux : nullptr;
        J->hits = out->element_hits;
        J->stats = out->stats;
        if (d->flags & ST_TRACE_RECORDS) {
            J->records = (float4*)out->records;
            J->rec_elem = out->record_elements;
        }
    }
    return ST_OK;
}

int validate_trace(st_scene s, const st_trace_desc* d, const st_outputs* out) {
    if (!s || !d) return fail(ST_ERR_INVALID, "null scene or trace descriptor");
    if (!s->sun_set) return fail(ST_ERR_STATE, "st_scene_set_sun must be called before tracing");
    if (d->num_rays == 0) return fail(ST_ERR_INVALID, "num_rays must be positive");
    if (d->ray_begin >= d->num_rays) return fail(ST_ERR_INVALID, "ray_begin beyond num_rays");
    uint64_t count = d->ray_count ? d->ray_count : d->num_rays - d->ray_begin;
    if (d->ray_begin + count > d->num_rays) return fail(ST_ERR_INVALID, "ray range exceeds num_rays");
    if (s->sun.position_sampling == ST_POS_HALTON && d->ray_begin + count > kHaltonMaxRays)
        return fail(ST_ERR_INVALID, "Halton positions take a 32-bit ray index: ray_begin + count must be <= 2^32");
    if (d->max_depth < 1 || d->max_depth > kMaxDepthLimit)
        return fail(ST_ERR_INVALID, "max_depth %d outside [1, %d]", d->max_depth, kMaxDepthLimit);
    if (d->block_size != 0 && d->block_size != 64 && d->block_size != 128 && d->block_size != 256)
        return fail(ST_ERR_INVALID, "block_size must be 0, 64, 128 or 256");
    if (d->grid_blocks < 0) return fail(ST_ERR_INVALID, "negative grid_blocks");
    if (d->waves_per_simd < 0) return fail(ST_ERR_INVALID, "negative waves_per_simd");
    if (d->kernel < 0 || d->kernel > ST_KERNEL_QUEUED_GROUPED)
        return fail(ST_ERR_INVALID, "unknown kernel %d", d->kernel);
    if ((d->flags & ST_TRACE_RECORDS) && (!out || !out->records))
        return fail(ST_ERR_INVALID, "ST_TRACE_RECORDS needs outputs->records");
    if (d->flux_nx > 0 && d->flux_ny > 0 && (!out || !out->flux))
        return fail(ST_ERR_INVALID, "flux grid requested without outputs->flux");
    return ST_OK;
}
}  // namespace st

namespace st {
// Launch shape of one st_trace / st_trace_noalloc call. An explicit form runs
// as asked. Under ST_KERNEL_AUTO the grouped form is used when the autotuner
// chose it for this scene (its choice stands for the lists it was tuned on),
// or when reflected rays mix listed segments with BVH ones (no or too few
// second-reflection lists).
int trace_shape(st_scene s, const st_trace_desc* d, const JobDev& J, LaunchShape* out, uint64_t* chunks) {
    LaunchShape sh = pick_shape(J.ray_count, d->block_size, d->grid_blocks, d->kernel, s->topo.depth,
                                s->tuned_valid ? &s->tuned : nullptr, s->n, d->waves_per_simd);
    if (d->waves_per_simd > 0 && !trace_variant_exists(sh.block, sh.stack, sh.waves, sh.queued))
        return fail(ST_ERR_INVALID, "no %s trace variant for block %d at %d waves/SIMD (BVH depth %d)",
                    sh.queued ? "queued" : "whole-ray", sh.block, d->waves_per_simd, s->topo.depth);
    bool group = false;
    if (d->kernel == ST_KERNEL_QUEUED_GROUPED) {
        group = true;
    } else if (d->kernel == ST_KERNEL_AUTO) {
        const bool tuned = s->tuned_valid && d->block_size <= 0 && d->grid_blocks <= 0 &&
                           s->tuned_list_build == s->list_builds;
        group = tuned ? s->tuned.group_bvh : mixed_bvh_segments(s, J);
    }
    use_grouped_bvh(sh, group);
    if (d->kernel == ST_KERNEL_QUEUED_GROUPED && !sh.group_bvh)
        return fail(ST_ERR_INVALID, "no grouped trace variant for block %d, stack %d, %d waves/SIMD", sh.block,
                    sh.stack, sh.waves);
    *chunks = (J.ray_count + sh.chunk - 1) / sh.chunk;
    if (*chunks >= (1ull << 32)) return fail(ST_ERR_INVALID, "too many rays for one call");
    *out = sh;
    return ST_OK;
}
}  // namespace st

extern "C" int st_trace(st_scene s, const st_trace_desc* d, const st_outputs* out) {
    ST_TRY(validate_trace(s, d, out));
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    if (!(d->flags & ST_TRACE_BVH_ONLY))
        ST_TRY(scene_ensure_lists(s, d->ray_count ? d->ray_count : d->num_rays - d->ray_begin, false));
    JobDev J;
    ST_TRY(fill_job(s, d, out, s->d_prims.ptr, s->d_nodes.ptr, s->d_sun.ptr, &J));
    set_trace_order(&J, s->sun.position_sampling);
    hipStream_t st = d->stream ? (hipStream_t)d->stream : s->ctx->stream;
    ST_TRY(order_before_trace(s->ctx, st));
    if ((d->flags & ST_TRACE_ZERO_OUTPUTS) && out) {
        if (J.flux) HIP_TRY(launch_fill_u64(st, J.flux, (uint64_t)d->flux_nx * d->flux_ny, 0), "zero flux");
        if (out->element_hits) HIP_TRY(launch_fill_u64(st, out->element_hits, (uint64_t)s->n, 0), "zero hits");
        if (out->stats) HIP_TRY(launch_fill_u64(st, out->stats, ST_STAT_COUNT, 0), "zero stats");
    }
    LaunchShape sh;
    uint64_t chunks = 0;
    ST_TRY(trace_shape(s, d, J, &sh, &chunks));
    if (launch_injected()) return fail(ST_ERR_INJECTED, "injected launch failure (trace)");
    // the chunk counter and the flux copies belong to the scene: a trace on
    // another stream than the scene's last one waits for that one's trace
    ST_TRY(order_scene_traces(s, st));
    int64_t* const flux_out = J.flux;
    uint32_t copies = 1;
    if (J.flux) ST_TRY(route_flux_copies(s->d_flux_copies, &J, 1, J.flux_bins, J.ray_count, st, &copies));
    HIP_TRY(launch_trace(st, nullptr, nullptr, 1, chunks, (uint32_t*)s->d_counter.ptr, sh, J), "trace kernel");
    if (copies > 1) HIP_TRY(launch_flux_gather(st, J.flux, copies, J.flux_bins, 1, flux_out), "flux gather");
    ST_TRY(mark_scene_trace(s, st));
    s->last_launch = launch_info(sh, chunks, copies);
    if (!d->stream) HIP_TRY(hipStreamSynchronize(st), "trace sync");
    return ST_OK;
}

// st_trace without device allocation: no lazy candidate-list build (lists the
// scene already holds are used), the chunk counter and the flux copies in the
// device's scratch instead of the scene's buffers. Both are cleared by kernels
// at each call (graph-safe); calls overlapping on one GPU share the scratch
// and race.
extern "C" int st_trace_noalloc(st_scene s, const st_trace_desc* d, const st_outputs* out) {
    ST_TRY(validate_trace(s, d, out));
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    char* const scr = (char*)scratch_of(s->ctx->device);
    if (!scr) return fail(ST_ERR_STATE, "no trace scratch on device %d (context not created)", s->ctx->device);
    JobDev J;
    ST_TRY(fill_job(s, d, out, s->d_prims.ptr, s->d_nodes.ptr, s->d_sun.ptr, &J));
    set_trace_order(&J, s->sun.position_sampling);
    hipStream_t st = d->stream ? (hipStream_t)d->stream : s->ctx->stream;
    ST_TRY(order_before_trace(s->ctx, st));
    LaunchShape sh;
    uint64_t chunks = 0;
    ST_TRY(trace_shape(s, d, J, &sh, &chunks));
    // flux copies as st_trace takes them, halved while they exceed the scratch
    uint32_t copies = J.flux ? flux_copies(J.flux_bins, J.ray_count) : 1;
    while (copies > 1 && (uint64_t)copies * J.flux_bins * sizeof(int64_t) > kScratchFluxBytes) copies >>= 1;
    if (copies > 1 && (uint64_t)copies * J.flux_bins * sizeof(int64_t) > kScratchFluxBytes)
        return fail(ST_ERR_INVALID, "flux copies of %llu bytes exceed the %zu-byte trace scratch",
                    (unsigned long long)((uint64_t)copies * J.flux_bins * sizeof(int64_t)), kScratchFluxBytes);
    if (launch_injected()) return fail(ST_ERR_INJECTED, "injected launch failure (trace)");
    if ((d->flags & ST_TRACE_ZERO_OUTPUTS) && out) {
        if (J.flux) HIP_TRY(launch_fill_u64(st, J.flux, (uint64_t)d->flux_nx * d->flux_ny, 0), "zero flux");
        if (out->element_hits) HIP_TRY(launch_fill_u64(st, out->element_hits, (uint64_t)s->n, 0), "zero hits");
        if (out->stats) HIP_TRY(launch_fill_u64(st, out->stats, ST_STAT_COUNT, 0), "zero stats");
    }
    uint32_t* const counter = (uint32_t*)scr;
    HIP_TRY(launch_fill_u64(st, counter, kScratchCounterBytes / sizeof(uint64_t), 0), "clear chunk counter");
    int64_t* const flux_out = J.flux;
    if (copies > 1) {
        int64_t* const region = (int64_t*)(scr + kScratchCounterBytes);
        HIP_TRY(launch_fill_u64(st, region, (uint64_t)copies * J.flux_bins, 0), "clear flux copies");
        J.flux = region;
        J.flux_copy_mask = copies - 1;
    }
    HIP_TRY(launch_trace(st, nullptr, nullptr, 1, chunks, counter, sh, J), "trace kernel");
    if (copies > 1) HIP_TRY(launch_flux_gather(st, J.flux, copies, J.flux_bins, 1, flux_out), "flux gather");
    s->last_launch = launch_info(sh, chunks, copies);
    if (!d->stream) HIP_TRY(hipStreamSynchronize(st), "trace sync");
    return ST_OK;
}

extern "C" int st_scene_last_launch(st_scene s, st_launch_info* info) {
    if (!s || !info) return fail(ST_ERR_INVALID, "null argument");
    *info = s->last_launch;
    return ST_OK;
}

extern "C" int st_scene_build_lists(st_scene s) {
    if (!s) return fail(ST_ERR_INVALID, "null scene");
    if (!s->sun_set) return fail(ST_ERR_STATE, "st_scene_set_sun must be called before building lists");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    return scene_ensure_lists(s, 0, true);
}

extern "C" int st_scene_list_info(st_scene s, uint64_t* grid_items, uint64_t* reflection_items, uint64_t* pair_items) {
    if (!s) return fail(ST_ERR_INVALID, "null scene");
    if (grid_items) *grid_items = s->grid_valid ? s->grid_items : 0;
    if (reflection_items) *reflection_items = s->bounce_valid ? s->bounce_items : 0;
    if (pair_items) *pair_items = s->pair_valid ? s->pair_items : 0;
    return ST_OK;
}

extern "C" int st_scene_autotune(st_scene s, uint64_t sample_rays, int32_t max_depth, st_tune_result* out) {
    if (!s || !out) return fail(ST_ERR_INVALID, "null argument");
    if (!s->sun_set) return fail(ST_ERR_STATE, "st_scene_set_sun must be called before tuning");
    if (sample_rays < 64 || sample_rays > (1ull << 32)) return fail(ST_ERR_INVALID, "sample_rays outside [64, 2^32]");
    if (max_depth < 1 || max_depth > kMaxDepthLimit) return fail(ST_ERR_INVALID, "bad max_depth %d", max_depth);
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    ST_TRY(scene_ensure_lists(s, sample_rays, true));  // tune the path production traces will take
    hipStream_t st = s->ctx->stream;
    DevBuf hits, stats, flux;
    ST_TRY(hits.ensure(sizeof(uint64_t) * s->n, "tune hits"));
    ST_TRY(stats.ensure(sizeof(uint64_t) * ST_STAT_COUNT, "tune stats"));
    ST_TRY(flux.ensure(sizeof(int64_t) * 256 * 128, "tune flux"));
    st_trace_desc d;
    memset(&d, 0, sizeof(d));
    d.num_rays = sample_rays;
    d.max_depth = max_depth;
    d.flux_element = -1;
    d.flux_nx = 256;
    d.flux_ny = 128;
    st_outputs o;
    memset(&o, 0, sizeof(o));
    o.flux = (int64_t*)flux.ptr;
    o.element_hits = (uint64_t*)hits.ptr;
    o.stats = (uint64_t*)stats.ptr;
    JobDev J;
    int32_t fslot = -1;
    ST_TRY(resolve_flux_element(s, -1, 256, 128, &fslot));
    if (fslot < 0) d.flux_nx = d.flux_ny = 0;
    ST_TRY(fill_job(s, &d, &o, s->d_prims.ptr, s->d_nodes.ptr, s->d_sun.ptr, &J));
    set_trace_order(&J, s->sun.position_sampling);
    uint32_t copies = 1;
    if (J.flux) ST_TRY(route_flux_copies(s->d_flux_copies, &J, 1, J.flux_bins, sample_rays, st, &copies));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "event");
    HIP_TRY(hipEventCreate(&e1), "event");
    auto time_shape = [&](const LaunchShape& sh, double* ms) -> int {
        const uint64_t chunks = (sample_rays + sh.chunk - 1) / sh.chunk;
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {  // first run warms caches and code
            HIP_TRY(hipEventRecord(e0, st), "event");
            HIP_TRY(launch_trace(st, nullptr, nullptr, 1, chunks, (uint32_t*)s->d_counter.ptr, sh, J), "trace");
            HIP_TRY(hipEventRecord(e1, st), "event");
            HIP_TRY(hipEventSynchronize(e1), "event sync");
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, e0, e1), "event time");
            if (rep > 0) best = std::min(best, t);
        }
        *ms = best;
        return ST_OK;
    };
    const int cus = device_cu_count();
    const LaunchShape def = pick_shape(sample_rays, 0, 0, ST_KERNEL_AUTO, s->topo.depth, nullptr, s->n, 0);
    double def_ms = 0;
    int rc = time_shape(def, &def_ms);
    LaunchShape best = def;
    double best_ms = def_ms;
    int trials = 1;
    const bool log_trials = getenv("ST_TUNE_LOG") != nullptr;  // one stderr line per trial
    if (log_trials)
        fprintf(stderr, "st_scene_autotune: default block %d stack %d waves %d queued %d grid %d chunk %d: %.3f ms\n",
                def.block, def.stack, def.waves, (int)def.queued, def.grid, def.chunk, def_ms);
    // grids: 0 = the blocks the device holds at once, else 4 / 8 / 16 x 256 threads per CU
    const int blocks[] = {64, 128, 256}, stacks[] = {0, 16, 32}, wavesv[] = {3, 4, 5, 6, 7, 8},
              per_cu[] = {0, 4, 8, 16};
    for (int rf = 0; rf < 2; ++rf)
    for (int b : blocks)
        for (int stk : stacks)
            for (int w : wavesv)
                for (int pc : per_cu) {
                    if (rc != ST_OK) break;
                    if (stack_entries(stk) < s->topo.depth || !trace_variant_exists(b, stk, w, rf != 0)) continue;
                    LaunchShape sh = def;
                    sh.block = b;
                    sh.stack = stk;
                    sh.waves = w;
                    sh.queued = rf != 0;
                    if (pc > 0) {
                        sh.grid = cus * pc * 256 / b;
                    } else {
                        const int rw = trace_resident_waves(sh);
                        if (rw <= 0) continue;
                        sh.grid = rw / (b / 64);
                    }
                    const uint64_t wv = (uint64_t)sh.grid * (uint64_t)(b / 64);
                    uint64_t chunk = (sample_rays / (4 * wv) + 63) / 64 * 64;
                    sh.chunk = (int)std::min<uint64_t>(std::max<uint64_t>(chunk, 64), 2048);
                    double ms = 0;
                    rc = time_shape(sh, &ms);
                    ++trials;
                    if (log_trials)
                        fprintf(stderr, "st_scene_autotune: block %d stack %d waves %d queued %d grid %d chunk %d: %.3f ms\n",
                                b, stk, w, rf, sh.grid, sh.chunk, ms);
                    if (rc == ST_OK && ms < 0.98 * best_ms) {
                        best_ms = ms;
                        best = sh;
                    }
                }
    // the queued winner with list and BVH segments traced in separate steps
    // (pays where many segments have no list, e.g. rays after two mirrors at
    // low sun): kept unless it is more than 1 % slower, as the form that holds
    // up when the sun moves lower (at high sun the two run within noise)
    if (rc == ST_OK && best.queued && !best.lds_hits && grouped_variant_exists(best.block, best.stack, best.waves)) {
        LaunchShape sh = best;
        sh.group_bvh = true;
        double ms = 0;
        rc = time_shape(sh, &ms);
        ++trials;
        if (log_trials)
            fprintf(stderr, "st_scene_autotune: block %d stack %d waves %d queued 2 grid %d chunk %d: %.3f ms\n",
                    sh.block, sh.stack, sh.waves, sh.grid, sh.chunk, ms);
        if (rc == ST_OK && ms < 1.01 * best_ms) {
            best_ms = ms;
            best = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // clear the copies (their sums go to the tuning buffer, discarded)
    if (copies > 1) HIP_TRY(launch_flux_gather(st, J.flux, copies, J.flux_bins, 1, o.flux), "flux gather");
    HIP_TRY(hipStreamSynchronize(st), "tune sync");
    if (rc != ST_OK) return rc;
    s->tuned = best;
    s->tuned_valid = true;
    s->tuned_list_build = s->list_builds;
    out->block = best.block;
    out->grid = best.grid;
    out->stack = best.stack;
    out->waves = best.waves;
    out->queued = best.queued ? (best.group_bvh ? 2 : 1) : 0;
    out->trials = trials;
    out->best_ms = best_ms;
    out->default_ms = def_ms;
    return ST_OK;
}

extern "C" int st_flux_to_watts(st_scene s, const st_trace_desc* d, const int64_t* flux_dev, double* flux_wm2_dev,
                                double* power_per_ray) {
    if (!s || !d) return fail(ST_ERR_INVALID, "null argument");
    if (!s->sun_set) return fail(ST_ERR_STATE, "sun not set");
    if (d->num_rays == 0) return fail(ST_ERR_INVALID, "num_rays must be positive");
    int32_t slot = -1;
    ST_TRY(resolve_flux_element(s, d->flux_element, d->flux_nx, d->flux_ny, &slot));
    int32_t fe = slot >= 0 ? s->topo.slot_elem[slot] : -1;
    double bin_area = 1.0;
    int32_t nbins = 0;
    if (fe >= 0 && flux_dev && flux_wm2_dev) {
        const st_element_f64& e = s->host_elems[fe];
        const double W = e.surface == ST_SURFACE_CYLINDER ? kPi * e.dim_x : e.dim_x;  // 2 pi r, r = w/2
        bin_area = (W / d->flux_nx) * (e.dim_y / d->flux_ny);
        nbins = d->flux_nx * d->flux_ny;
    }
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    hipStream_t st = d->stream ? (hipStream_t)d->stream : s->ctx->stream;
    ST_TRY(order_before_trace(s->ctx, st));
    double* ppr_dev = nullptr;
    DevBuf tmp;
    if (power_per_ray) {
        ST_TRY(tmp.ensure(sizeof(double), "power per ray"));
        ppr_dev = (double*)tmp.ptr;
    }
    HIP_TRY(launch_flux_watts(st, flux_dev, nbins, (const SunFrame*)s->d_sun.ptr, (double)d->num_rays, bin_area,
                              nbins ? flux_wm2_dev : nullptr, ppr_dev),
            "flux normalisation");
    if (power_per_ray) HIP_TRY(hipMemcpyAsync(power_per_ray, ppr_dev, sizeof(double), hipMemcpyDeviceToHost, st), "download");
    HIP_TRY(hipStreamSynchronize(st), "sync");
    return ST_OK;
}

// ---------------------------------------------------------------------------
// batched sweep
namespace st {
static bool same_batch(const BatchCache& c, const st_batch_desc* d) {
    if (c.steps != d->steps || c.max_depth != d->max_depth || c.seed != d->seed || c.flags != d->flags ||
        c.flux_element != d->flux_element || c.flux_nx != d->flux_nx || c.flux_ny != d->flux_ny ||
        c.block != d->block_size || c.grid != d->grid_blocks || c.stream != d->stream)
        return false;
    if (memcmp(&c.sun, &d->sun, sizeof(st_sun_desc)) != 0) return false;
    if (memcmp(c.target, d->target, sizeof(c.target)) != 0) return false;
    if (c.suns.size() != (size_t)d->steps * 3 || memcmp(c.suns.data(), d->sun_vectors, sizeof(double) * 3 * d->steps) != 0)
        return false;
    if (c.budgets.size() != (size_t)d->steps || memcmp(c.budgets.data(), d->budgets, sizeof(uint64_t) * d->steps) != 0)
        return false;
    return true;
}
}  // namespace st

// per step: FP64 re-aim (+ frames for the reflection lists), refit, sun plane
static int batch_scene_states(st_scene s, hipStream_t st) {
    BatchCache& b = s->batch;
    const int32_t K = b.steps, n = s->n;
    HIP_TRY(launch_frames_batch(st, (const st_element_f64*)s->d_elems.ptr, n, (const int32_t*)s->d_elem_slot.ptr, K,
                                (const double*)b.d_sun_units.ptr, (const double*)b.d_target.ptr,
                                (double*)b.d_ref_aabb.ptr, (float*)b.d_bounds.ptr, (float4*)b.d_prims.ptr,
                                b.lists ? (double*)b.d_frames.ptr : nullptr),
            "batched frames");
    HIP_TRY(launch_refit(st, (const int32_t*)s->d_child.ptr, (const int32_t*)s->d_level_nodes.ptr,
                         s->level_offsets.data(), (int32_t)s->level_offsets.size() - 1, s->n_nodes, n, K,
                         (const float*)b.d_bounds.ptr, (float4*)b.d_nodes.ptr),
            "batched refit");
    HIP_TRY(launch_sun_plane(st, (const double*)b.d_ref_aabb.ptr, n, K, (const double*)b.d_sun_units.ptr, &b.sun,
                             (SunFrame*)b.d_sun.ptr),
            "batched sun plane");
    return ST_OK;
}

// Sizes the per-step grids and reflection lists of a new batch. The grid
// item capacity per step is the longest total over the steps whose grid pays,
// found by one eager pass over the scene states (the runs recompute the same
// states, so the same steps get grids); steps over it trace the BVH.
static int batch_plan_lists(st_scene s, hipStream_t st) {
    BatchCache& b = s->batch;
    const int32_t K = b.steps, n = s->n;
    b.grids = b.lists = false;
    b.grid_cap = 0;
    if (b.flags & ST_TRACE_BVH_ONLY) return ST_OK;
    const uint64_t list_tasks = (uint64_t)K * n;
    const uint64_t list_bytes = list_tasks * (sizeof(int32_t) * kMaxBounceList + sizeof(uint2) + 9 * sizeof(double));
    b.lists = list_bytes <= kBatchListBytes / 2 && list_tasks * kMaxBounceList < (1ull << 32);
    if (b.lists) {
        ST_TRY(b.d_frames.ensure(sizeof(double) * 9 * list_tasks, "batch frames"));
        ST_TRY(b.d_bounce_range.ensure(sizeof(uint2) * list_tasks, "batch reflection ranges"));
        ST_TRY(b.d_bounce_scratch.ensure(sizeof(int32_t) * kMaxBounceList * list_tasks, "batch reflection lists"));
        ST_TRY(b.d_sup_spheres.ensure(sizeof(double4) * (size_t)n, "batch element spheres"));
        ST_TRY(b.d_sup_range.ensure(sizeof(uint2) * (size_t)n, "batch re-aim list ranges"));
        ST_TRY(b.d_sup_scratch.ensure(sizeof(int32_t) * kMaxBounceList * (size_t)n, "batch re-aim lists"));
    }
    const uint64_t grid_bytes = (uint64_t)K * (3 * kGridCells + 1) * sizeof(uint32_t);
    if (grid_bytes > kBatchListBytes / 4 || n < kGridMinElements) return ST_OK;
    ST_TRY(b.d_grid_counts.ensure(sizeof(uint32_t) * kGridCells * (size_t)K, "batch grid counts"));
    ST_TRY(b.d_grid_cursor.ensure(sizeof(uint32_t) * kGridCells * (size_t)K, "batch grid cursor"));
    ST_TRY(b.d_grid_start.ensure(sizeof(uint32_t) * (kGridCells + 1) * (size_t)K, "batch grid start"));
    ST_TRY(b.d_grid_stats.ensure(sizeof(uint32_t) * 2 * (size_t)K, "batch grid stats"));
    ST_TRY(batch_scene_states(s, st));
    HIP_TRY(launch_fill_u64(st, b.d_grid_counts.ptr, (uint64_t)K * kGridCells / 2, 0), "batch grid clear");
    HIP_TRY(launch_grid_count(st, (const float*)b.d_bounds.ptr, (const float4*)b.d_prims.ptr,
                              (const uint32_t*)s->d_meta.ptr, n, K, (const SunFrame*)b.d_sun.ptr, sun_cone_tan(b.sun),
                              (uint32_t*)b.d_grid_counts.ptr, (uint32_t*)b.d_grid_start.ptr,
                              (uint32_t*)b.d_grid_stats.ptr, nullptr, nullptr, 0),
            "batch grid count");
    std::vector<uint32_t> gs(2 * (size_t)K);
    HIP_TRY(hipMemcpyAsync(gs.data(), b.d_grid_stats.ptr, sizeof(uint32_t) * gs.size(), hipMemcpyDeviceToHost, st),
            "batch grid stats");
    HIP_TRY(hipStreamSynchronize(st), "batch grid sync");
    uint64_t cap = 0;
    for (int32_t k = 0; k < K; ++k)
        if (grid_lists_pay(gs[2 * k], gs[2 * k + 1])) cap = std::max<uint64_t>(cap, gs[2 * k]);
    if (cap == 0 || cap * K * sizeof(int32_t) > kBatchListBytes / 4) return ST_OK;
    ST_TRY(b.d_grid_items.ensure(sizeof(int32_t) * cap * (size_t)K, "batch grid items"));
    b.grid_cap = (uint32_t)cap;
    b.grids = true;
    return ST_OK;
}

static int batch_enqueue(st_scene s, hipStream_t st) {
    BatchCache& b = s->batch;
    const int32_t K = b.steps, n = s->n;
    if (b.flux_out) HIP_TRY(launch_fill_u64(st, b.flux_out, (uint64_t)K * b.nbins, 0), "zero flux");
    if (b.hits_out) HIP_TRY(launch_fill_u64(st, b.hits_out, (uint64_t)K * n, 0), "zero hits");
    if (b.stats_out) HIP_TRY(launch_fill_u64(st, b.stats_out, (uint64_t)K * ST_STAT_COUNT, 0), "zero stats");
    ST_TRY(batch_scene_states(s, st));
    if (b.grids) {
        HIP_TRY(launch_fill_u64(st, b.d_grid_counts.ptr, (uint64_t)K * kGridCells / 2, 0), "batched grid clear");
        HIP_TRY(launch_grid_count(st, (const float*)b.d_bounds.ptr, (const float4*)b.d_prims.ptr,
                                  (const uint32_t*)s->d_meta.ptr, n, K, (const SunFrame*)b.d_sun.ptr,
                                  sun_cone_tan(b.sun), (uint32_t*)b.d_grid_counts.ptr, (uint32_t*)b.d_grid_start.ptr,
                                  (uint32_t*)b.d_grid_stats.ptr, (JobDev*)b.d_jobs.ptr,
                                  (const int32_t*)b.d_grid_items.ptr, b.grid_cap),
                "batched grid count");
        HIP_TRY(launch_fill_u64(st, b.d_grid_cursor.ptr, (uint64_t)K * kGridCells / 2, 0), "batched grid cursor clear");
        HIP_TRY(launch_grid_fill(st, (const float*)b.d_bounds.ptr, (const float4*)b.d_prims.ptr,
                                 (const uint32_t*)s->d_meta.ptr, n, K, (const SunFrame*)b.d_sun.ptr,
                                 sun_cone_tan(b.sun), (const uint32_t*)b.d_grid_start.ptr,
                                 (uint32_t*)b.d_grid_cursor.ptr, (int32_t*)b.d_grid_items.ptr, b.grid_cap),
                "batched grid fill");
    }
    if (b.lists) {
        // every mirror is re-aimed at the target: candidates over all re-aims
        // once, then each step keeps those its cones reach
        HIP_TRY(launch_reaim_lists(st, (const st_element_f64*)s->d_elems.ptr, n, (const int32_t*)s->d_elem_slot.ptr,
                                   (const uint32_t*)s->d_meta.ptr, (const double*)b.d_target.ptr,
                                   sun_cone_angle(b.sun), (double4*)b.d_sup_spheres.ptr, (uint2*)b.d_sup_range.ptr,
                                   (int32_t*)b.d_sup_scratch.ptr),
                "re-aim candidate lists");
        HIP_TRY(launch_cone_lists(st, (const st_element_f64*)s->d_elems.ptr, K * n, n, n, s->n_nodes,
                                  (const int32_t*)s->d_elem_slot.ptr, (const uint32_t*)s->d_meta.ptr,
                                  (const double*)b.d_frames.ptr, (const float*)b.d_bounds.ptr,
                                  (const float4*)b.d_nodes.ptr, (const SunFrame*)b.d_sun.ptr, sun_cone_angle(b.sun),
                                  nullptr, nullptr, nullptr, (uint2*)b.d_bounce_range.ptr,
                                  (int32_t*)b.d_bounce_scratch.ptr, (const uint2*)b.d_sup_range.ptr,
                                  (const int32_t*)b.d_sup_scratch.ptr),
                "batched reflection lists");
    }
    JobDev dummy;
    memset(&dummy, 0, sizeof(dummy));
    HIP_TRY(launch_trace(st, (const JobDev*)b.d_jobs.ptr, (const uint64_t*)b.d_prefix.ptr, K, b.total_chunks,
                         (uint32_t*)b.d_counter.ptr, b.shape, dummy),
            "batched trace");
    if (b.flux_copies > 1)
        HIP_TRY(launch_flux_gather(st, (int64_t*)b.d_flux_copies.ptr, b.flux_copies, b.nbins, K, b.flux_out),
                "batched flux gather");
    return ST_OK;
}

extern "C" int st_trace_batch(st_scene s, const st_batch_desc* d, int64_t* flux, uint64_t* element_hits,
                              uint64_t* stats) {
    if (!s || !d) return fail(ST_ERR_INVALID, "null argument");
    if (d->steps <= 0 || d->steps > 65535) return fail(ST_ERR_INVALID, "steps %d outside [1, 65535]", d->steps);
    if (!d->sun_vectors || !d->budgets) return fail(ST_ERR_INVALID, "null sun vectors or budgets");
    if (d->max_depth < 1 || d->max_depth > kMaxDepthLimit) return fail(ST_ERR_INVALID, "bad max_depth %d", d->max_depth);
    if (d->block_size != 0 && d->block_size != 64 && d->block_size != 128 && d->block_size != 256)
        return fail(ST_ERR_INVALID, "block_size must be 0, 64, 128 or 256");
    if (d->flags & ST_TRACE_RECORDS) return fail(ST_ERR_INVALID, "records are not available in batched sweeps");
    if (d->sun.shape != ST_SUN_PILLBOX && d->sun.shape != ST_SUN_GAUSSIAN) return fail(ST_ERR_INVALID, "bad sun shape");
    if (!(d->sun.param >= 0) || d->sun.param > 0.2) return fail(ST_ERR_INVALID, "bad sun angle");
    if (d->sun.position_sampling < ST_POS_PHILOX || d->sun.position_sampling > ST_POS_STRATIFIED)
        return fail(ST_ERR_INVALID, "bad position sampling");
    for (int k = 0; k < d->steps; ++k) {
        const double* v = d->sun_vectors + 3 * k;
        if (!(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 0)) return fail(ST_ERR_INVALID, "step %d: zero sun vector", k);
        if (d->budgets[k] == 0) return fail(ST_ERR_INVALID, "step %d: zero ray budget", k);
        if (d->sun.position_sampling == ST_POS_HALTON && d->budgets[k] > kHaltonMaxRays)
            return fail(ST_ERR_INVALID, "step %d: Halton positions take a 32-bit ray index (budget > 2^32)", k);
    }
    int32_t fslot = -1;
    ST_TRY(resolve_flux_element(s, d->flux_element, d->flux_nx, d->flux_ny, &fslot));
    if (fslot >= 0 && !flux) return fail(ST_ERR_INVALID, "flux grid requested without a flux buffer");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    hipStream_t st = d->stream ? (hipStream_t)d->stream : s->ctx->stream;
    ST_TRY(order_before_trace(s->ctx, st));
    BatchCache& b = s->batch;
    const bool reuse = b.valid && same_batch(b, d) && b.flux_out == flux && b.hits_out == element_hits &&
                       b.stats_out == stats;
    if (!reuse) {
        if (s->batch_graph_exec) {
            (void)hipGraphExecDestroy(s->batch_graph_exec);
            s->batch_graph_exec = nullptr;
        }
        b.valid = false;
        const int32_t K = d->steps, n = s->n;
        b.steps = K;
        b.max_depth = d->max_depth;
        b.seed = d->seed;
        b.flags = d->flags;
        b.flux_element = d->flux_element;
        b.flux_nx = d->flux_nx;
        b.flux_ny = d->flux_ny;
        b.block = d->block_size;
        b.grid = d->grid_blocks;
        b.stream = d->stream;
        b.sun = d->sun;
        if (!(b.sun.dni > 0)) b.sun.dni = 1000.0;
        memcpy(b.target, d->target, sizeof(b.target));
        b.suns.assign(d->sun_vectors, d->sun_vectors + 3 * (size_t)K);
        b.budgets.assign(d->budgets, d->budgets + K);
        b.flux_out = flux;
        b.hits_out = element_hits;
        b.stats_out = stats;
        b.nbins = fslot >= 0 ? (uint64_t)d->flux_nx * d->flux_ny : 0;
        std::vector<double> units(3 * (size_t)K);
        for (int k = 0; k < K; ++k) {
            const double* v = d->sun_vectors + 3 * k;
            const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            for (int j = 0; j < 3; ++j) units[3 * k + j] = v[j] / l;
        }
        ST_TRY(b.d_sun_units.ensure(sizeof(double) * 3 * K, "batch suns"));
        ST_TRY(b.d_target.ensure(sizeof(double) * 3, "batch target"));
        ST_TRY(b.d_ref_aabb.ensure(sizeof(double) * 6 * (size_t)n * K, "batch AABBs"));
        ST_TRY(b.d_bounds.ensure(sizeof(float) * 6 * (size_t)n * K, "batch bounds"));
        ST_TRY(b.d_prims.ensure(sizeof(float4) * 4 * (size_t)n * K, "batch primitives"));
        ST_TRY(b.d_nodes.ensure(sizeof(float4) * 4 * (size_t)s->n_nodes * K, "batch nodes"));
        ST_TRY(b.d_sun.ensure(sizeof(SunFrame) * K, "batch sun frames"));
        ST_TRY(b.d_jobs.ensure(sizeof(JobDev) * K, "batch jobs"));
        ST_TRY(b.d_prefix.ensure(sizeof(uint64_t) * (K + 1), "batch chunk prefix"));
        ST_TRY(b.d_counter.ensure(sizeof(uint32_t) * 2, "batch counter"));
        HIP_TRY(hipMemcpyAsync(b.d_sun_units.ptr, units.data(), sizeof(double) * 3 * K, hipMemcpyHostToDevice, st), "upload");
        HIP_TRY(hipMemcpyAsync(b.d_target.ptr, d->target, sizeof(double) * 3, hipMemcpyHostToDevice, st), "upload");
        ST_TRY(batch_plan_lists(s, st));
        uint64_t total_rays = 0;
        for (int k = 0; k < K; ++k) total_rays += d->budgets[k];
        b.shape = pick_shape(total_rays, d->block_size, d->grid_blocks, ST_KERNEL_AUTO, s->topo.depth,
                             s->tuned_valid ? &s->tuned : nullptr, s->n, 0);
        use_grouped_bvh(b.shape, b.lists);  // per-step first-reflection lists, no pair lists
        std::vector<JobDev> jobs(K);
        std::vector<uint64_t> prefix(K + 1, 0);
        for (int k = 0; k < K; ++k) {
            st_trace_desc td;
            memset(&td, 0, sizeof(td));
            td.num_rays = d->budgets[k];
            td.seed = d->seed;
            td.job_id = (uint32_t)k;
            td.max_depth = d->max_depth;
            td.flags = d->flags;
            td.flux_element = d->flux_element;
            td.flux_nx = d->flux_nx;
            td.flux_ny = d->flux_ny;
            st_outputs o;
            memset(&o, 0, sizeof(o));
            o.flux = flux ? flux + (size_t)k * b.nbins : nullptr;
            o.element_hits = element_hits ? element_hits + (size_t)k * n : nullptr;
            o.stats = stats ? stats + (size_t)k * ST_STAT_COUNT : nullptr;
            ST_TRY(fill_job(s, &td, &o, (const float4*)b.d_prims.ptr + (size_t)k * n * 4,
                            (const float4*)b.d_nodes.ptr + (size_t)k * s->n_nodes * 4,
                            (const SunFrame*)b.d_sun.ptr + k, &jobs[k]));
            set_trace_order(&jobs[k], b.sun.position_sampling);
            // grid pointers are set by each run's grid scan; lists of step k
            // are indexed by element within the step's range block
            jobs[k].grid_start = nullptr;
            jobs[k].grid_items = nullptr;
            if (b.lists) {
                jobs[k].bounce_range = (const uint2*)b.d_bounce_range.ptr + (size_t)k * n;
                jobs[k].bounce_items = (const int32_t*)b.d_bounce_scratch.ptr;
            }
            prefix[k + 1] = prefix[k] + (d->budgets[k] + b.shape.chunk - 1) / b.shape.chunk;
        }
        b.total_chunks = prefix[K];
        if (b.total_chunks >= (1ull << 32)) return fail(ST_ERR_INVALID, "too many rays for one sweep");
        b.flux_copies = 1;
        if (b.nbins)
            ST_TRY(route_flux_copies(b.d_flux_copies, jobs.data(), K, b.nbins, total_rays, st, &b.flux_copies));
        HIP_TRY(hipMemcpyAsync(b.d_jobs.ptr, jobs.data(), sizeof(JobDev) * K, hipMemcpyHostToDevice, st), "upload");
        HIP_TRY(hipMemcpyAsync(b.d_prefix.ptr, prefix.data(), sizeof(uint64_t) * (K + 1), hipMemcpyHostToDevice, st),
                "upload");
        HIP_TRY(hipMemsetAsync(b.d_counter.ptr, 0, sizeof(uint32_t) * 2, st), "counter clear");
        HIP_TRY(hipStreamSynchronize(st), "batch setup sync");
        b.valid = true;
        if (d->use_graph) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
            int rc = batch_enqueue(s, st);
            hipError_t ce = hipStreamEndCapture(st, &g);
            if (rc != ST_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ce != hipSuccess) return hip_fail(ce, "end capture");
            hipError_t ie = hipGraphInstantiate(&s->batch_graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                s->batch_graph_exec = nullptr;
                return hip_fail(ie, "graph instantiate");
            }
        }
    }
    if (d->use_graph && s->batch_graph_exec) {
        HIP_TRY(hipGraphLaunch(s->batch_graph_exec, st), "graph launch");
    } else {
        ST_TRY(batch_enqueue(s, st));
    }
    if (!d->stream) HIP_TRY(hipStreamSynchronize(st), "batch sync");
    return ST_OK;
}

extern "C" int st_calibrate_valu(st_context ctx, double* lane_ops_per_s) {
    if (!ctx || !lane_ops_per_s) return fail(ST_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const int blocks = device_cu_count() * 16;
    DevBuf out;
    ST_TRY(out.ensure(sizeof(float) * 256 * (size_t)blocks, "calibration output"));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "event");
    HIP_TRY(hipEventCreate(&e1), "event");
    double ops = 0.0;
    float best = 1e30f;
    int rc = ST_OK;
    for (int rep = 0; rep < 5 && rc == ST_OK; ++rep) {  // the first launch warms up
        hipError_t e = hipEventRecord(e0, ctx->stream);
        if (e == hipSuccess) e = launch_valu_calibrate(ctx->stream, (float*)out.ptr, blocks, &ops);
        if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) rc = hip_fail(e, "VALU calibration");
        else if (rep > 0) best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != ST_OK) return rc;
    *lane_ops_per_s = ops / (best * 1e-3);
    return ST_OK;
}

extern "C" int st_debug_inject_fault(int32_t what, int32_t count) {
    if (count < 0) return fail(ST_ERR_INVALID, "negative count");
    if (what == 0) g_fail_alloc.store(count);
    else if (what == 1) g_fail_launch.store(count);
    else return fail(ST_ERR_INVALID, "unknown fault kind %d", what);
    return ST_OK;
}
