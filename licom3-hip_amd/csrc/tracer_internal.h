// tracer_internal.h -- context and kernel interfaces shared by the HIP
// sources of liblicom_tracer.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "licom_tracer.h"

namespace licom {

// Static grid and coefficients as seen by the kernels (device pointers).
struct KGrid {
    int imt, jmt, km;
    const int32_t *kmt, *kmu;
    const double *tarea, *tarea_r, *htw, *hts, *hun, *hue;
    const double *ohbt, *ohbu, *dtn, *dts, *dte, *dtw, *ahf;
    const double *au0, *aus, *auw, *ausw;
    const double *dzp, *odzp, *odzt, *pen, *to, *so, *eosc;
    const double *dxu, *dyu;   // U-cell widths (flux scheme faces; zeros otherwise)
};

struct KPar {
    double dts, onbc, oncc, dwndmix, gamma, od0, od0cp, ah;
    int hmix, boundary_restore, sssnorm;
    int adv;    // LICOM_ADV_CENTERED / _FLUX / _TSPAS
    int nss;    // divisor of the centered scheme's ax/ay/az increments
};

struct Scratch {
    double *stf_pre = nullptr;   // [jmt][imt] 0.5*h0f*onbc + 0.5*h0l
    double *work_u = nullptr;    // [jmt][imt] T->U average of stf_pre
    double *stf_n = nullptr;     // [2][jmt][imt] surface flux seen by tracer n
    double *temp11 = nullptr;    // [jmt][imt] SSSNORM integrand
    double *partials = nullptr;  // [2*nblk] block partial sums (hi, lo)
    double *err_norm2 = nullptr; // [1]
    double *uw = nullptr, *vs = nullptr;   // [km][jmt][imt] face velocities
    uint8_t *flags = nullptr;    // [2][km][jmt][imt] tspas limiter flags
    double *d2tk = nullptr;      // [2][km][jmt][imt] del4 first pass
    double *ori = nullptr;       // [2][km][jmt][imt] explicit update
    int *mix_count = nullptr;    // [2] columns convadj mixed in the last step,
                                 // tracer columns whose new state is not finite
    int nblk = 0;
};

struct TimedEvent {
    const char *name;
    hipEvent_t start, stop;
};

}  // namespace licom

struct licom_ctx {
    int nx, ny, km, imt, jmt;
    size_t n2, n3;
    licom_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    // launch shape of the single-model column solve (solve_rows/solve_cols):
    // wide grids take one row of 128 columns per workgroup, so that the
    // lines a row segment shares with its neighbours (the S half of each
    // 3-D array sits 64 bytes off the T half at odd n3/8) are mostly written
    // and read by one workgroup
    int solve_shape = 0;
    // licomqice [jmt][imt] (device) written by steps with LICOM_STEP_ICESNOW
    double *qice = nullptr;
    // tspas limiter flags [2][km][jmt][imt] (device, 1 byte) written by the
    // fused FP64 and mixed k-marches when set (licom_tracer_set_flag_output)
    uint8_t *flags_out = nullptr;
    licom::KGrid g{};
    licom::KPar kp{};
    double area_t = 0.0;
    // device allocations owned by the context
    std::vector<void *> allocs;
    size_t scratch_bytes = 0;
    licom::Scratch s;
    // host copies of the derived fields (for licom_grid_export)
    std::vector<int32_t> h_kmt, h_kmu;
    std::vector<double> h_tarea_r, h_ohbt, h_ohbu, h_dtn, h_dts, h_dte, h_dtw, h_ahf;
    std::vector<double> h_dzp, h_odzp, h_odzt;
    // profiling
    bool profiling = false;
    std::vector<licom::TimedEvent> pending;
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<int> prof_counts;
};

namespace licom {

void set_error(const char *fmt, ...);
int hip_status(hipError_t e, const char *what);
// tracer_api.hip: argument checks and path choice of an FP64 step
int plan_step(licom_ctx *c, const licom_fields *f, int &flags, bool &split, const char *who);
// tracer_noalloc.hip: device address of the no-allocation step's static
// scratch, looked up once per device (called by licom_tracer_create with
// the device current)
int noalloc_init(int device);

// Kernel launchers.  All enqueue on ctx->stream.
// tracer_kernels.hip: one kernel per reference stage (cross-check path)
int launch_step_simple(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags);
int launch_convadj(licom_ctx *ctx, const licom_fields *f, const licom_diag *d);
int launch_icesnow(licom_ctx *ctx, double *at, double *qice, int reset);
// shared pieces of tracer_kernels.hip used by both paths
int launch_surface(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, const Scratch &s);
int launch_halo(licom_ctx *ctx, const licom_fields *f, int flags);
// tracer_fused.hip: prelude / k-march / column solve (production path)
// s: the step's scratch (the context's, or the no-allocation entry point's)
int launch_step_fused(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags,
                      const Scratch &s);
// tracer_mixed.hip: the mixed-precision form
int launch_step_mixed(licom_ctx *ctx, const licom_fields_mixed *f, const licom_diag_mixed *d,
                      int flags);

int launch_stream_copy(void *dst, const void *src, size_t bytes, hipStream_t st);

// XCD-aware 1-D launch of a gx x gy grid of column blocks.  The hardware
// deals the workgroups of a launch to the eight XCDs round-robin, so
// workgroup b runs on XCD b % 8; xcd_block gives it block L = (b % 8) * per +
// b / 8 (row-major), so each XCD walks a contiguous run of blocks and the
// cache lines neighbouring blocks share (halo columns, lines that straddle a
// block edge) meet in that XCD's L2.  Launch xcd_blocks() workgroups; those
// with L >= gx * gy return at once.
struct XcdGrid {
    int gx, n, per;
};
inline XcdGrid xcd_grid(unsigned gx, unsigned gy) {
    const int n = (int)(gx * gy);
    return XcdGrid{(int)gx, n, (n + 7) / 8};
}
inline unsigned xcd_blocks(const XcdGrid &x) { return 8u * (unsigned)x.per; }
__device__ __forceinline__ bool xcd_block(const XcdGrid &x, int &bx, int &by) {
    const int b = (int)blockIdx.x, L = (b & 7) * x.per + (b >> 3);
    if (L >= x.n) return false;
    by = L / x.gx;
    bx = L - by * x.gx;
    return true;
}

// Raw buffer access to one plane (up to 2 GiB).  A lane whose access is
// masked off gets an offset past the end of the buffer: the hardware then
// returns 0 for its load and drops its store, so no branch surrounds the
// instruction.  Branch-free loads and stores keep the compiler's count of
// outstanding vector-memory operations exact, so a wait for an earlier load
// does not also wait for later stores (on gfx9 one vmcnt covers both).
typedef int licom_v2i __attribute__((ext_vector_type(2)));
constexpr uint32_t kBufOff = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t buf_off(bool ok, size_t idx, uint32_t esize) {
    return ok ? (uint32_t)idx * esize : kBufOff;
}
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(licom_v2i, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, off, 0, 0);
}
// s_waitcnt vmcnt(0) that the compiler's own wait insertion knows about
__device__ __forceinline__ void wait_all_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// RAII HIP-event bracket used when ctx->profiling is set.
struct Timer {
    licom_ctx *ctx;
    const char *name;
    hipEvent_t a{}, b{};
    Timer(licom_ctx *c, const char *n);
    ~Timer();
};

}  // namespace licom
