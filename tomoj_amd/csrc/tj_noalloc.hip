// tj_noalloc.hip -- projectAndDifference on caller-owned device buffers with no
// device allocation: tj_project_and_difference_noalloc.
//
// tj_project_and_difference keeps its workspaces (projection table, ray norms,
// error partials, row flags, per-image errors) in its context and allocates them
// on first use.  This variant takes the volume, the measured images and the
// difference images as device pointers from the caller and places every
// workspace in one file-scope device array, `scratch`, of TJ_NOALLOC_SCRATCH_BYTES
// per device.  The array's device address is looked up once per device with
// hipGetSymbolAddress and only that address is handed to kernels and copies.
// Calls on one device are serialised by a per-device mutex; the array keeps the
// last call's contents, and every workspace entry a call reads is written by
// that call first (nothing accumulates across calls).
#include "tj_device.h"
#include "../../include/tomoj_hip.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

namespace tjk {
hipError_t norm_table(hipStream_t, const TjGeomDev&, const TjProj*, int, float*);
hipError_t fp_residual(hipStream_t, const TjGeomDev&, const float*, const TjProj*, const int*, int, int, const float*,
                       int, int, const float*, double, float*, double*, uint8_t*, uint8_t*, double*,
                       hipEvent_t mid = nullptr, int gb = 1, int* clear = nullptr, bool pf = true,
                       float* proj = nullptr, float* norm = nullptr);
hipError_t plane_flags(hipStream_t, const float*, size_t, int, int, uint8_t*);
void set_max_lds(int);
}  // namespace tjk

void tj_geom_to_dev(const tj_geometry* geo, TjGeomDev* out);
int tj_fp_start_row(const tj_geometry& g, int start_y);

static __device__ __align__(256) unsigned char scratch[TJ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_dev_mu[kMaxDevices];
unsigned char* g_scratch[kMaxDevices];  // device address of `scratch` per device (null: not looked up yet)

int nfail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    tj_record_error(buf);
    return code;
}

#define TJ_NA_HIP(call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return nfail(TJ_EDEVICE, "tj_project_and_difference_noalloc: %s: %s", #call,                 \
                         hipGetErrorString(e_));                                                         \
    } while (0)

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the workspaces inside the scratch array for n images.
struct Layout {
    size_t pp, slot, norm, rerr, rbad, rowflag, err, pflag, total;
};

Layout layout(const TjGeomDev& g, int n)
{
    const size_t ntile = (size_t)(g.pw + TJ_BLOCK - 1) / TJ_BLOCK, nrb = (size_t)(g.ph + 3) / 4;
    Layout L{};
    size_t o = 0;
    L.pp = o;
    o = align256(o + sizeof(TjProj) * n);
    L.slot = o;
    o = align256(o + sizeof(int) * n);
    L.norm = o;
    o = align256(o + sizeof(float) * n * (size_t)g.pw);
    L.rerr = o;
    o = align256(o + sizeof(double) * n * nrb * ntile);
    L.rbad = o;
    o = align256(o + (size_t)n * g.ph * ntile);
    L.rowflag = o;
    o = align256(o + (size_t)n * g.ph);
    L.err = o;
    o = align256(o + sizeof(double) * n);
    L.pflag = o;
    o = align256(o + (size_t)g.vh + 1);
    L.total = o;
    return L;
}

}  // namespace

extern "C" size_t tj_noalloc_scratch_bytes(const tj_geometry* geo, int n)
{
    if (!geo || n <= 0 || geo->vol_w <= 0 || geo->vol_h <= 0 || geo->proj_w <= 0 || geo->proj_h <= 0) return 0;
    TjGeomDev g;
    tj_geom_to_dev(geo, &g);
    return layout(g, n).total;
}

extern "C" int tj_project_and_difference_noalloc(int device, const tj_geometry* geo, const float* vol_yzx,
                                                 const float* measured, const double* euler, int n, double factor,
                                                 int start_y, int end_y, float* diff, double* errors,
                                                 void* stream)
{
    static const char* who = "tj_project_and_difference_noalloc";
    if (!geo || !vol_yzx || !measured || !euler || !diff || !errors) return nfail(TJ_EINVAL, "%s: null argument", who);
    if (n <= 0) return nfail(TJ_EINVAL, "%s: %d images", who, n);
    if (geo->vol_w <= 0 || geo->vol_h <= 0 || geo->vol_d <= 0 || geo->proj_w <= 0 || geo->proj_h <= 0)
        return nfail(TJ_EINVAL, "%s: non-positive size", who);
    if (geo->storage != TJ_STORAGE_F32) return nfail(TJ_EUNSUP, "%s: storage type %d not supported", who, geo->storage);
    if (start_y < 0 || end_y > geo->vol_h || start_y > end_y)
        return nfail(TJ_EINVAL, "%s: row range [%d,%d) outside the volume height %d", who, start_y, end_y, geo->vol_h);
    if (geo->long_object_comp)
        return nfail(TJ_EUNSUP, "%s: long-object compensation needs the context path (tj_project_and_difference)", who);
    TjGeomDev g;
    tj_geom_to_dev(geo, &g);
    const Layout L = layout(g, n);
    if (L.total > (size_t)TJ_NOALLOC_SCRATCH_BYTES)
        return nfail(TJ_ENOMEM, "%s: %d images of %dx%d need %zu scratch bytes, the maximum is %zu", who, n, g.pw,
                     g.ph, L.total, (size_t)TJ_NOALLOC_SCRATCH_BYTES);

    // The one forward kernel this entry point runs is the LDS kernel fused with
    // the residual; its preconditions are those tj_project_and_difference checks
    // before it picks the LDS kernel: one vertical tilt axis for every image,
    // scale 1, whole volume rows, equal row geometry.
    const int start = tj_fp_start_row(*geo, start_y);
    const TjProj a = tj_make_proj(euler);
    for (int k = 1; k < n; ++k) {
        const TjProj b = tj_make_proj(euler + 9 * k);
        if (!b.axis || b.e[4] != 1.0 || b.icf != a.icf || b.zcf != a.zcf || b.divy != a.divy || b.diry != a.diry)
            return nfail(TJ_EUNSUP, "%s: image %d does not share image 0's single vertical tilt axis geometry", who, k);
    }
    const int zero = 0;
    if (!tj_fp_lds_ok(g, &a, &zero, 1, start))
        return nfail(TJ_EUNSUP, "%s: geometry outside the single-axis LDS path (scale 1, vertical axis, whole rows)", who);

    int ndev = 0;
    TJ_NA_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= kMaxDevices)
        return nfail(TJ_EINVAL, "%s: device %d of %d", who, device, ndev);
    std::lock_guard<std::mutex> lock(g_dev_mu[device]);
    TJ_NA_HIP(hipSetDevice(device));
    if (!g_scratch[device]) {
        void* p = nullptr;
        TJ_NA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return nfail(TJ_EDEVICE, "%s: no device address for the scratch array", who);
        int max_lds = 65536;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess)
            tjk::set_max_lds(max_lds);
        g_scratch[device] = (unsigned char*)p;
    }
    unsigned char* base = g_scratch[device];
    hipStream_t st = (hipStream_t)stream;
    TjProj* d_pp = (TjProj*)(base + L.pp);
    int* d_slot = (int*)(base + L.slot);
    float* d_norm = (float*)(base + L.norm);
    double* d_rerr = (double*)(base + L.rerr);
    uint8_t* d_rbad = base + L.rbad;
    uint8_t* d_rowflag = base + L.rowflag;
    double* d_err = (double*)(base + L.err);
    uint8_t* d_pflag = base + L.pflag;

    // projection table and slot list, through small host chunks (each copy
    // completes before its chunk is refilled)
    {
        constexpr int CH = 64;
        TjProj hp[CH];
        int hs[CH];
        for (int k0 = 0; k0 < n; k0 += CH) {
            const int m = (n - k0 < CH) ? n - k0 : CH;
            for (int k = 0; k < m; ++k) {
                hp[k] = tj_make_proj(euler + 9 * (size_t)(k0 + k));
                hs[k] = k0 + k;
            }
            TJ_NA_HIP(hipMemcpyAsync(d_pp + k0, hp, sizeof(TjProj) * m, hipMemcpyHostToDevice, st));
            TJ_NA_HIP(hipMemcpyAsync(d_slot + k0, hs, sizeof(int) * m, hipMemcpyHostToDevice, st));
            TJ_NA_HIP(hipStreamSynchronize(st));
        }
    }
    // the LDS forward kernel needs a finite volume (non-finite samples take the
    // global-load kernel in the context path)
    const size_t plane = (size_t)g.vd * g.vw;
    TJ_NA_HIP(tjk::plane_flags(st, vol_yzx, plane, 0, g.vh, d_pflag));
    {
        uint8_t hf[4096];
        for (int y0 = 0; y0 < g.vh; y0 += (int)sizeof hf) {
            const int m = (g.vh - y0 < (int)sizeof hf) ? g.vh - y0 : (int)sizeof hf;
            TJ_NA_HIP(hipMemcpyAsync(hf, d_pflag + y0, m, hipMemcpyDeviceToHost, st));
            TJ_NA_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < m; ++k)
                if (hf[k]) return nfail(TJ_EUNSUP, "%s: volume row %d holds non-finite values", who, y0 + k);
        }
    }
    TJ_NA_HIP(tjk::norm_table(st, g, d_pp, n, d_norm));
    TJ_NA_HIP(tjk::fp_residual(st, g, vol_yzx, d_pp, d_slot, n, start, d_norm, 4, 4, measured, factor, diff, d_rerr,
                               d_rbad, d_rowflag, d_err, nullptr, 0));
    TJ_NA_HIP(hipMemcpyAsync(errors, d_err, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    TJ_NA_HIP(hipStreamSynchronize(st));
    return TJ_OK;
}
