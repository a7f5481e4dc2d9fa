// tracer_noalloc.hip -- licom_tracer_step_noalloc: the FP64 step for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops), with the step's scratch in a file-scope device array
// instead of the context's allocation.
//
// The array (kNoallocBytes) holds the scratch of grids up to
// 32 * imt*jmt*km + 44 * imt*jmt bytes (about 4.7 M cells at km = 55, e.g.
// 360 x 218 x 55): the face velocities u_wface/v_sface, the explicit update
// of both tracers, the surface and SSSNORM fields and the mixing counter.
// Each device the library runs on holds one copy until the process ends.
// Every call on a device shares that device's copy: calls that overlap on
// one device (several streams or threads) race.  Each call clears the part
// it uses with a fill kernel before the step, so no state carries over.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <type_traits>

#include "tracer_internal.h"

namespace licom {
namespace {

constexpr size_t kNoallocBytes = size_t(160) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};       // device address of `scratch`, per device
int g_status[kMaxDevices] = {};      // 0: not looked up, 1: ok, < 0: failed

__global__ __launch_bounds__(256) void k_zero16(uint4 *p, size_t n) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = make_uint4(0u, 0u, 0u, 0u);
}

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// Carves the scratch of one step out of base (NULL: sizes only); returns
// the bytes used.
size_t carve(const licom_ctx *c, unsigned char *base, Scratch &s) {
    const size_t n2 = c->n2, n3 = c->n3;
    size_t off = 0;
    auto take = [&](auto **dst, size_t bytes) {
        *dst = base ? reinterpret_cast<std::remove_pointer_t<decltype(dst)>>(base + off) : nullptr;
        off += al256(bytes);
    };
    take(&s.stf_pre, n2 * 8);
    take(&s.work_u, n2 * 8);
    take(&s.stf_n, 2 * n2 * 8);
    take(&s.temp11, n2 * 8);
    take(&s.partials, (size_t)2 * c->s.nblk * 8);
    take(&s.err_norm2, 8);
    take(&s.uw, n3 * 8);
    take(&s.vs, n3 * 8);
    take(&s.ori, 2 * n3 * 8);
    take(&s.mix_count, 8);
    s.flags = nullptr;
    s.d2tk = nullptr;
    s.nblk = c->s.nblk;
    return off;
}

}  // namespace

int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return LICOM_EINVAL;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_status[device] != 0) return g_status[device] > 0 ? LICOM_OK : g_status[device];
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_status[device] = LICOM_EHIP;
        return LICOM_EHIP;
    }
    g_ptr[device] = p;
    g_status[device] = 1;
    return LICOM_OK;
}

}  // namespace licom

using namespace licom;

extern "C" int licom_tracer_step_noalloc(licom_ctx *c, const licom_fields *f, const licom_diag *d,
                                         int flags) {
    bool split = false;
    if (int rc = plan_step(c, f, flags, split, "licom_tracer_step_noalloc")) return rc;
    if (flags & LICOM_STEP_STAGED) {
        set_error("licom_tracer_step_noalloc runs the fused kernels only: it needs an even imt, "
                  "16-byte aligned at/atb/ws/akt, no LICOM_STEP_STAGED and not ICESNOW with "
                  "VTL_SCRATCH");
        return LICOM_ENOTSUP;
    }
    Scratch s;
    const size_t need = carve(c, nullptr, s);
    if (need > kNoallocBytes) {
        set_error("licom_tracer_step_noalloc: this grid needs %zu bytes of scratch, above the "
                  "%zu-byte static buffer (32*imt*jmt*km + 44*imt*jmt bytes must fit)",
                  need, kNoallocBytes);
        return LICOM_EINVAL;
    }
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device < 0 || c->device >= kMaxDevices || g_status[c->device] <= 0) {
            set_error("licom_tracer_step_noalloc: the static scratch of device %d was not set up "
                      "(hipGetSymbolAddress failed at licom_tracer_create)", c->device);
            return LICOM_EHIP;
        }
        base = g_ptr[c->device];
    }
    carve(c, static_cast<unsigned char *>(base), s);
    {
        Timer t(c, "k_zero_scratch");
        const size_t n16 = need / 16;
        k_zero16<<<(unsigned)((n16 + 255) / 256), 256, 0, c->stream>>>(static_cast<uint4 *>(base), n16);
    }
    if (int rc = hip_status(hipGetLastError(), "scratch clear launch")) return rc;
    const int run = split ? flags & ~LICOM_STEP_CONVADJ : flags;
    if (int rc = launch_step_fused(c, f, d, run, s)) return rc;
    return split ? launch_convadj(c, f, d) : LICOM_OK;
}
