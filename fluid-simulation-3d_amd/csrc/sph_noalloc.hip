// sph_noalloc.hip -- sph_update_noalloc: one Update(dt) of one scene held in
// caller-owned device memory, with every intermediate in a fixed device array
// of this file (`scratch`), so the call allocates nothing.
//
// No counterpart in the reference.  It serves callers that cannot allocate
// (serving loops and harnesses that forbid hipMalloc): the same kernels as
// sph_create + sph_upload + sph_update + sph_download under the cell-grid key,
// with the handle's buffers replaced by regions of `scratch`.
//
// `scratch` is one file-scope device array per GPU, held until the process
// ends: SPH_NOALLOC_SCRATCH_BYTES (24,645,376, about 23.5 MiB) for up to kNoAllocMaxParticles
// particles and kNoAllocMaxBuckets buckets.  Its device address is looked up
// once per device with hipGetSymbolAddress (the array's name in host code is
// a host-side stand-in, not a device address) and kept per device.  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU (from two
// threads, or on two streams without ordering) race and give wrong results.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>

#include "../../include/sph_amd.h"
#include "sph_engine.h"
#include "sph_launch.h"

namespace sph {
int set_last_error(int code, const std::string &msg);  // sph_capi.cpp
}

namespace {

constexpr uint32_t kNoAllocMaxParticles = SPH_NOALLOC_MAX_PARTICLES;
// the cell-grid key sizes a scene of n <= 2^16 particles to at most
// max(8n, 2^20) = 2^20 buckets (grid_table_size)
constexpr uint32_t kNoAllocMaxBuckets = 1u << 20;
constexpr uint32_t kCountChunk = sph::kCountChunkBuckets;  // buckets per workgroup of the count scan

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Byte offsets of the regions of `scratch`, for the largest scene.
struct Layout {
    size_t ed, sd, scan_err, sflags, chain, bounds, count;
    size_t pos, vel, xps, vs, vs2, dens, ps;
    size_t keys0, keys1, ids, sids, ncnt, order, nbr;
    size_t total;
};

constexpr Layout make_layout()
{
    const size_t N = kNoAllocMaxParticles, T = kNoAllocMaxBuckets;
    const size_t chain_words = (T + kCountChunk - 1) / kCountChunk + 1;
    Layout L{};
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.ed = take(sizeof(EngineDesc));
    L.sd = take(sizeof(SceneDesc));
    L.scan_err = take(sizeof(uint32_t));
    L.sflags = take(sizeof(uint32_t));
    L.chain = take(sizeof(uint64_t) * chain_words);
    L.bounds = take(sizeof(uint32_t) * (T + 1));
    L.count = take(sizeof(uint32_t) * (T + 1));  // the buckets, then one scene flag
    L.pos = take(16 * N);
    L.vel = take(16 * N);
    L.xps = take(16 * N);
    L.vs = take(16 * N);
    L.vs2 = take(16 * N);
    L.dens = take(16 * N);
    L.ps = take(16 * N);
    L.keys0 = take(4 * N);
    L.keys1 = take(4 * N);
    L.ids = take(4 * N);
    L.sids = take(4 * N);
    L.ncnt = take(4 * N);
    L.order = take(8 * N);
    L.nbr = take(4 * 27 * N);  // neighbour-mask records: up to 18 words per particle used
    L.total = o;
    return L;
}

constexpr Layout kLayout = make_layout();
static_assert(kLayout.total == SPH_NOALLOC_SCRATCH_BYTES, "the scratch size stated in sph_amd.h");

static __device__ __align__(256) unsigned char scratch[kLayout.total];

// The descriptors of the call, written into `scratch` by a kernel from its
// arguments (no host-to-device copy).
__global__ void k_write_descs(const EngineDesc e, const SceneDesc d, EngineDesc *__restrict__ ed,
                              SceneDesc *__restrict__ sd)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *ed = e;
        *sd = d;
    }
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_base(unsigned char **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return sph::set_last_error(SPH_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return sph::set_last_error(SPH_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return sph::set_last_error(SPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                        hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch[dev];
    return SPH_OK;
}

template <typename T>
T *at(unsigned char *base, size_t off)
{
    return reinterpret_cast<T *>(base + off);
}

int hip_fail(hipError_t e, const char *what)
{
    return sph::set_last_error(SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" int sph_update_noalloc(const sph_params *p, uint32_t n, float dt, const float *pos4_in,
                                  const float *vel4_in, float *pos4_out, float *vel4_out,
                                  float *dens2_out, void *stream, int wait)
{
    if (!p || !pos4_in || !vel4_in || !pos4_out || !vel4_out)
        return sph::set_last_error(SPH_ERR_INVALID, "null argument");
    if (n == 0 || n > kNoAllocMaxParticles)
        return sph::set_last_error(SPH_ERR_INVALID,
                                   "sph_update_noalloc: n must be in [1, SPH_NOALLOC_MAX_PARTICLES = 65536]");
    if (!isfinite(dt)) return sph::set_last_error(SPH_ERR_INVALID, "dt must be finite");
    if (const char *m = sph::check_params(*p)) return sph::set_last_error(SPH_ERR_INVALID, m);
    const uint32_t T = sph::grid_table_size(*p, n);
    if (T == 0 || T > kNoAllocMaxBuckets)
        return sph::set_last_error(SPH_ERR_INVALID, "sph_update_noalloc: bucket count beyond its maximum");
    unsigned char *base = nullptr;
    if (int rc = scratch_base(&base)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);

    EngineDesc E{};
    E.n_total = n;
    E.table_total = T;
    E.n_scenes = 1;
    E.scene_shift = 32;
    E.morton_drop = 0;
    E.scan_spin = sph::Engine::kScanSpinDefault;
    SceneDesc D{};
    sph::fill_scene_desc(*p, n, T, true, &D);
    D.p_off = 0;
    D.key_off = 0;
    D.index = 0;
    D.dt = dt;

    // The handle's buffers, as regions of `scratch`.  Stage A and the sort
    // read the caller's input in place (particle order: resident slot k holds
    // particle k); the viscosity pass writes the new state into `scratch`,
    // sorted, and it is unpermuted into the caller's output.
    const uint32_t cb = (T + kCountChunk - 1) / kCountChunk;
    sph::DevState s{};
    s.ed = at<EngineDesc>(base, kLayout.ed);
    s.sd = at<SceneDesc>(base, kLayout.sd);
    s.scene_of = nullptr;  // one scene
    s.multi = false;
    s.grid = true;
    s.pos = const_cast<float4 *>(reinterpret_cast<const float4 *>(pos4_in));  // read only
    s.vel = const_cast<float4 *>(reinterpret_cast<const float4 *>(vel4_in));  // read only
    s.ids = at<uint32_t>(base, kLayout.ids);
    s.keys[0] = at<uint32_t>(base, kLayout.keys0);
    s.keys[1] = at<uint32_t>(base, kLayout.keys1);
    s.order = at<uint2>(base, kLayout.order);
    s.count = at<uint32_t>(base, kLayout.count);
    s.count_size = T + 1;
    s.chain = at<uint64_t>(base, kLayout.chain);
    s.count_blocks = cb;
    s.xps = at<float4>(base, kLayout.xps);
    s.vs = at<float4>(base, kLayout.vs);
    s.vs2 = at<float4>(base, kLayout.vs2);
    s.dens = at<float4>(base, kLayout.dens);
    s.ps = at<float4>(base, kLayout.ps);
    s.sids = at<uint32_t>(base, kLayout.sids);
    s.bounds = at<uint32_t>(base, kLayout.bounds);
    s.sflags = at<uint32_t>(base, kLayout.sflags);
    s.nbr = at<uint32_t>(base, kLayout.nbr);
    s.ncnt = at<uint32_t>(base, kLayout.ncnt);
    s.scan_err = at<uint32_t>(base, kLayout.scan_err);
    s.capacity = n;
    s.table_capacity = T;
    s.radix_blocks = 1;
    sph::DevState out = s;  // the viscosity pass and the unpermute: state in `scratch`
    out.pos = at<float4>(base, kLayout.pos);
    out.vel = at<float4>(base, kLayout.vel);
    const sph::LaunchCfg c = sph::default_launch_cfg(true);

    hipError_t e = hipSuccess;
#define NA(x, what)                                  \
    if ((e = (x)) != hipSuccess) return hip_fail(e, what)
    k_write_descs<<<1, 64, 0, st>>>(E, D, at<EngineDesc>(base, kLayout.ed), at<SceneDesc>(base, kLayout.sd));
    NA(hipGetLastError(), "descriptors");
    // the scan's chunk status words must start at zero (the reorder of every
    // call clears them again; this covers a call cut short before it)
    NA(sph::launch_clear_words(reinterpret_cast<uint32_t *>(s.chain), 2 * (cb + 1), st), "clear chain");
    NA(sph::launch_iota(s.ids, n, 0, st), "resident order");
    NA(sph::launch_predict(s, c, st), "predict");  // clears the counts first
    NA(sph::launch_bucket_sort(s, c, st), "bucket sort");
    NA(sph::launch_density(s, c, st), "density");
    NA(sph::launch_pressure(s, c, st), "pressure");
    NA(sph::launch_viscosity_integrate(out, c, st), "viscosity + integrate");
    NA(sph::launch_unpermute_state(out, 0, n, reinterpret_cast<float4 *>(pos4_out),
                                   reinterpret_cast<float4 *>(vel4_out), st),
       "unpermute state");
    if (dens2_out) NA(sph::launch_unpermute_dens(out, reinterpret_cast<float2 *>(dens2_out), st), "densities");
    if (!wait) return SPH_OK;
    NA(hipStreamSynchronize(st), "sync");
    uint32_t flag = 0;
    NA(hipMemcpy(&flag, s.scan_err, sizeof(flag), hipMemcpyDeviceToHost), "read device status");
    if (flag) {
        NA(sph::launch_clear_words(s.scan_err, 1, st), "clear device status");
        NA(hipStreamSynchronize(st), "sync");
        return sph::set_last_error(SPH_ERR_DEVICE,
                                   "the count scan's bounded look-back wait gave up: the outputs of "
                                   "this call (or of an earlier call that did not wait) are invalid");
    }
    return SPH_OK;
#undef NA
}
