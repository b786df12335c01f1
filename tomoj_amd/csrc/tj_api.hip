// tj_api.hip -- host side of libtomoj_hip.so: the C ABI declared in
// include/tomoj_hip.h, the projector context, and the C++ restatement of
// TomoJ's update loops (TomoReconstruction2.OSSART / osartIteration /
// orderShuffleART / WBP) that drive the kernels in tj_kernels.hip.
//
// Reference paths are relative to src/TomoJ_Base/src/main/java/fr/curie/tomoj/.
#include "tj_device.h"
#include "../../include/tomoj_hip.h"

#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

namespace tjk {
hipError_t fp(hipStream_t, const TjGeomDev&, const float*, const TjProj*, const int*, int, int, int,
              const uint8_t*, const float*, float*, float*, int, int, int, int gb = 1, bool pf = true);
hipError_t norm_table(hipStream_t, const TjGeomDev&, const TjProj*, int, float*);
void set_max_lds(int);
hipError_t norm_max_fill(hipStream_t, int, int, float*);
hipError_t residual(hipStream_t, int, int, int, const float*, const int*, const float*, const float*, double,
                    float*, double*, uint8_t*, double*);
hipError_t row_flags(hipStream_t, int, int, int, const float*, const int*, uint8_t*);
hipError_t fp_residual(hipStream_t, const TjGeomDev&, const float*, const TjProj*, const int*, int, int, const float*,
                       int, int, const float*, double, float*, double*, uint8_t*, uint8_t*, double*,
                       hipEvent_t mid = nullptr, int gb = 1, int* clear = nullptr, bool pf = true,
                       float* proj = nullptr, float* norm = nullptr);
hipError_t bp(hipStream_t, const TjGeomDev&, float*, const TjProj*, const int*, int, const float*, const int*,
              const uint8_t*, TjBpRow*, int, int, int, int, int, int, int, double2*, hipEvent_t mid = nullptr,
              int* slow = nullptr, int fast = 0, int slow_cleared = 0);
hipError_t zyx_to_yzx(hipStream_t, const float*, float*, int, int, int, int);
hipError_t yzx_to_zyx(hipStream_t, const float*, float*, int, int, int, int);
hipError_t plane_flags(hipStream_t, const float*, size_t, int, int, uint8_t*);
hipError_t phantom_project(hipStream_t, const double*, int, const TjProj*, int, int, int, double, double, double*);
hipError_t add_noise(hipStream_t, const double*, size_t, uint64_t, double, double*, int, float*);
hipError_t copy(hipStream_t, const void*, void*, size_t);
hipError_t square_diffs(hipStream_t, float*, size_t);
hipError_t fista(hipStream_t, float*, const float*, size_t, double);
hipError_t tv_denoise(hipStream_t, float*, float*, float*, int, int, int, float, float, float, int, uint8_t*);
hipError_t bayes_conv(hipStream_t, const float*, float*, int, int, int, double);
hipError_t median_cross(hipStream_t, const float*, float*, int, int, int);
hipError_t stack_stats(hipStream_t, const float*, size_t, double*, int*);
hipError_t collapse(hipStream_t, float*, size_t, double, double, float, int);
hipError_t bayes_end(hipStream_t, float*, float*, size_t, int, float, int, uint8_t*);
hipError_t normalize(hipStream_t, float*, int, int, int, const int*, double*);
hipError_t threshold(hipStream_t, float*, int, int, int, double, double, int, double*, uint8_t*, int*);
hipError_t roi_stats(hipStream_t, const float*, int, int, int, const int*, double*);
hipError_t poly_stats(hipStream_t, const float*, int, int, int, const int*, double*);
hipError_t affine(hipStream_t, const float*, float*, int, int, int, const double*, const int*, const double*, int,
                  int);
hipError_t weighting_init(hipStream_t, int, int, double, const double*, const double*, int, float*);
hipError_t ramp_init(hipStream_t, int, int, int, double, double, float*);
hipError_t weighting_apply(hipStream_t, int, int, int, const float*, const int*, const float*, float*, void*, size_t);
size_t weighting_workspace(int, int, int);
size_t fsc_workspace(int, int, int);
size_t ssnr_workspace(int, int);
hipError_t ssnr(hipStream_t, const float*, const float*, const float*, const float*, int, int, int, int, void*,
                size_t, double*, float*);
hipError_t fsc(hipStream_t, const float*, const float*, int, int, int, void*, size_t, double*);
}  // namespace tjk

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// The job engine (tj_batch.hip) reports its failures through tj_last_error too.
void tj_record_error(const char* msg) { g_last_error = msg; }

#define TJ_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(TJ_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                     \
    } while (0)

template <typename T>
static hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

template <typename T>
static void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct tj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    tj_geometry geo{};
    TjGeomDev g{};
    size_t npx = 0, plane = 0, nvox = 0;

    float* vol = nullptr;       // [vh][vd][vw]
    uint8_t* pflag = nullptr;   // [vh+1] non-finite flag per volume plane
    bool any_nonfinite = false; // host copy: some plane flag is set
    float* norm_tab = nullptr;  // [nproj][pw] ray norms of single-axis projections

    int nproj = 0;              // resident projections
    float* exps = nullptr;      // [nproj][ph][pw]
    TjProj* pp = nullptr;       // [nproj]
    std::vector<TjProj> hpp;
    int* order = nullptr;       // [nproj] device projection order for tj_run_ossart

    int cap = 0;                // slot capacity
    float *sproj = nullptr, *snorm = nullptr, *sdiff = nullptr, *swgt = nullptr;
    int swgt_cap = 0;
    double* row_err = nullptr;  // [cap][ph]
    uint8_t* rowflag = nullptr; // [cap][ph]
    uint8_t* rbad = nullptr;    // [cap][ph][column tiles] (fused forward + residual)
    TjBpRow* bprows = nullptr;  // [cap][vh]
    double* err = nullptr;      // [max(cap,nproj)]
    int* slot_proj = nullptr;   // [cap]
    std::vector<int> uploaded;  // the indices slot_proj holds (upload_added)
    int* iota = nullptr;        // [max(cap,nproj)] 0,1,2,...
    int iota_n = 0;
    double* h_err = nullptr;    // pinned [max(cap,nproj)]
    int h_err_n = 0;

    std::vector<int> added;
    bool have_diff = false;
    bool have_proj = false;

    // weighting filter
    int weight_kind = 0;        // 0 none, 1 Radermacher, 2 ramp
    float* mask = nullptr;      // [ph][pw]
    void* wspace = nullptr;
    size_t wspace_size = 0;

    // ReconstructionParameters.getAvailableIndexes: the images the fused loops
    // walk (empty: all)
    std::vector<int> avail;

    // FISTA wrapper (FistaProjector3D): volume at the start of the iteration, t
    int fista = 0;
    double fista_t = 0.0;
    float* prev = nullptr;      // [vh][vd][vw]

    // TVMProjector3D end-of-iteration denoising (TVMProjector3D.java:26-32)
    int tvm_iter = 0;           // tvmIteration; 0 = hook off
    float tvm_theta = 25.0f, tvm_g = 1.0f, tvm_dt = 0.1f;
    float* tv_f = nullptr;      // [vh][vd][vw] the volume entering the denoising
    float* tv_p = nullptr;      // [3][vh][vd][vw] px, py, pz

    // BayesianVoxelProjector3D (BayesianVoxelProjector3D.java:25-30): its
    // backprojection goes into the distance volume, folded into the volume at
    // the end of the iteration
    int bayes = 0;
    double bayes_alpha = 0.1, bayes_lambda = 0.1;
    float* dist = nullptr;      // [vh][vd][vw]

    // BgARTVoxelProjector3D (BgARTVoxelProjector3D.java:13-31)
    int bgart = 0, bg_dark = 1, bg_median = 1;
    double bg_k = 1.0, bg_value = NAN;
    float* bg_tmp = nullptr;    // [vh][vd][vw] copy read by the cross median
    double* bg_mm = nullptr;    // [2 x 4096] min/max partials
    int* bg_hist = nullptr;     // [256]

    // VoxelProjector3D.errorVolume (createErrorVolume, :858-864): while it
    // exists, every backprojection of differences also backprojects their
    // squares into it
    float* errvol = nullptr;    // [vh][vd][vw]

    // FP64 pair images of the backprojected images (LDS backprojection)
    double2* bpairs = nullptr;  // [bpairs_cap][ph][pw]
    int bpairs_cap = 0;

    // kernel selection and shapes (tj_set_tuning)
    tj_tuning tune{0, 2, 4, 5, 0, 4, 8, 3, 1, 0, 0};
    int* bpslow = nullptr;      // [2]: k_bp_rows' flag (a row the fast backprojection cannot take),
                                // and the count of launches the fast backprojection took

    // progress / cancellation
    std::atomic<double> completion{0.0};
    std::atomic<int> interrupted{0};

    // timing
    int timing_on = 0;
    tj_timing last_timing{};
    std::vector<hipEvent_t> events;
};

static void ctx_free(tj_ctx* c)
{
    dfree(c->bpslow);
    dfree(c->errvol);
    dfree(c->bpairs);
    dfree(c->prev);
    dfree(c->tv_f);
    dfree(c->tv_p);
    dfree(c->dist);
    dfree(c->bg_tmp);
    dfree(c->bg_mm);
    dfree(c->bg_hist);
    dfree(c->vol);
    dfree(c->pflag);
    dfree(c->norm_tab);
    dfree(c->exps);
    dfree(c->pp);
    dfree(c->order);
    dfree(c->sproj);
    dfree(c->snorm);
    dfree(c->sdiff);
    dfree(c->swgt);
    dfree(c->row_err);
    dfree(c->rowflag);
    dfree(c->bprows);
    dfree(c->err);
    dfree(c->slot_proj);
    dfree(c->iota);
    dfree(c->mask);
    if (c->wspace) (void)hipFree(c->wspace);
    c->wspace = nullptr;
    if (c->h_err) (void)hipHostFree(c->h_err);
    c->h_err = nullptr;
    for (auto e : c->events) (void)hipEventDestroy(e);
    c->events.clear();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

// Per-projection constants of projectJosephPartial (VoxelProjector3D.java:682-694),
// same operation order as the Java source (host code is built without FMA
// contraction).
TjProj tj_make_proj(const double* r)
{
    TjProj p{};
    for (int k = 0; k < 9; ++k) p.e[k] = r[k];
    p.icf = r[3] / r[0];
    p.zcf = (r[5] - (r[3] * r[2] / r[0]));
    p.divy = r[4] - (r[3] * r[1] / r[0]);
    p.yfx = r[1];
    p.zfx = r[2];
    p.divx = r[0];
    p.diry = -(r[5] - r[3] * r[2] / r[0]) / (r[4] - r[3] * r[1] / r[0]);
    p.dirx = (r[1] * (-p.diry) - r[2]) / r[0];
    p.norminc = std::sqrt(p.dirx * p.dirx + p.diry * p.diry + 1);
    p.axis = (r[1] == 0.0 && r[3] == 0.0 && r[5] == 0.0 && r[0] != 0.0 && r[4] != 0.0) ? 1 : 0;
    return p;
}

static bool all_axis(const tj_ctx* c, const std::vector<int>& idx)
{
    for (int i : idx)
        if (!c->hpp[i].axis) return false;
    return true;
}

// The LDS forward projection stages R consecutive volume rows for R consecutive
// detector rows and shares the row geometry between all projections of a
// launch.  Checked here with the kernel's own formulas: every projection must
// give the same y for each detector row, and the inside rows must map to
// consecutive volume rows.
bool tj_fp_lds_ok(const TjGeomDev& g, const TjProj* hpp, const int* idx, int n, int start_row)
{
    if (g.scale != 1.0 || n <= 0) return false;
    const TjProj& a = hpp[idx[0]];
    for (int k = 0; k < n; ++k) {
        const TjProj& b = hpp[idx[k]];
        if (!b.axis || b.e[4] != 1.0 || b.icf != a.icf || b.zcf != a.zcf || b.divy != a.divy || b.diry != a.diry)
            return false;
    }
    long first_yb = -1, first_j = -1;
    for (int j = std::max(start_row, 0); j < g.ph; ++j) {
        const int sy = start_row + ((j - start_row) & ~1);
        const double jinc = 1.0 / g.scale;
        double jc = (sy - 1.0 - g.pc1) / g.scale;
        jc += jinc;
        if (j != sy) jc += jinc;
        const double oldy = ((jc - a.icf * 0.0) - a.zcf * (-1 - g.vcz)) / a.divy;
        const double y = (oldy + a.diry) + g.vcy;
        if (!(y + TJ_ACCURACY >= 0 && y < g.bymax)) continue;
        const long yb = (y < 0) ? 0 : (long)y;
        if ((y < 0 ? 0.0f : (float)y - (float)yb) != 0.0f) return false;  // 4-tap rows
        if (first_j < 0) {
            first_j = j;
            first_yb = yb;
        } else if (yb != first_yb + (j - first_j)) {
            return false;
        }
    }
    return true;
}

static bool fp_lds_ok(const tj_ctx* c, const std::vector<int>& idx, int start_row)
{
    if (idx.empty() || c->any_nonfinite || !c->norm_tab) return false;
    return tj_fp_lds_ok(c->g, c->hpp.data(), idx.data(), (int)idx.size(), start_row);
}

// The LDS backprojection bounds the detector columns a block reaches by
// 256*|e00| + KC*|e02| and drops the multiplication by scale.
static bool bp_lds_ok(const tj_ctx* c, const std::vector<int>& idx)
{
    if (c->g.scale != 1.0) return false;
    for (int i : idx) {
        const TjProj& b = c->hpp[i];
        if (!b.axis || std::fabs(b.e[0]) > 1.0 || std::fabs(b.e[2]) > 1.0) return false;
    }
    return true;
}

static int pick_fp_kernel(const tj_ctx* c, const std::vector<int>& idx, int start_row)
{
    if (!all_axis(c, idx)) return 1;
    const int want = c->tune.fp_kernel;
    if (want == 1 || want == 2) return want;
    return fp_lds_ok(c, idx, start_row) ? 3 : 2;
}

static int pick_bp_kernel(const tj_ctx* c, const std::vector<int>& idx)
{
    if (!all_axis(c, idx)) return 1;
    const int want = c->tune.bp_kernel;
    if (want == 1 || want == 2) return want;
    return bp_lds_ok(c, idx) ? 3 : 2;
}

// Room for the pair images of n backprojected images (grows, never shrinks).
static int ensure_bpairs(tj_ctx* c, int n)
{
    if (n <= c->bpairs_cap) return TJ_OK;
    dfree(c->bpairs);
    c->bpairs_cap = 0;
    TJ_HIP(dalloc(&c->bpairs, (size_t)n * c->npx));
    c->bpairs_cap = n;
    return TJ_OK;
}

static int ensure_iota(tj_ctx* c, int n)
{
    if (n <= c->iota_n) return TJ_OK;
    dfree(c->iota);
    dfree(c->err);
    if (c->h_err) (void)hipHostFree(c->h_err);
    c->h_err = nullptr;
    std::vector<int> h(n);
    for (int k = 0; k < n; ++k) h[k] = k;
    TJ_HIP(dalloc(&c->iota, n));
    TJ_HIP(dalloc(&c->err, n));
    TJ_HIP(hipHostMalloc((void**)&c->h_err, sizeof(double) * n, hipHostMallocDefault));
    TJ_HIP(hipMemcpy(c->iota, h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    c->iota_n = n;
    c->h_err_n = n;
    return TJ_OK;
}

static int ensure_slots(tj_ctx* c, int n)
{
    if (n <= c->cap) return ensure_iota(c, n);
    dfree(c->sproj);
    dfree(c->snorm);
    dfree(c->sdiff);
    dfree(c->swgt);
    dfree(c->row_err);
    dfree(c->rowflag);
    dfree(c->rbad);
    dfree(c->bprows);
    dfree(c->slot_proj);
    dfree(c->bpslow);
    c->cap = 0;
    const size_t npx = c->npx;
    const size_t ntile = (size_t)(c->g.pw + 255) / 256;  // column tiles of the fused residual
    if (dalloc(&c->sproj, npx * n) || dalloc(&c->snorm, npx * n) || dalloc(&c->sdiff, npx * n) ||
        dalloc(&c->row_err, (size_t)c->g.ph * n * ntile) || dalloc(&c->rowflag, (size_t)c->g.ph * n) ||
        dalloc(&c->rbad, (size_t)c->g.ph * n * ntile) ||
        dalloc(&c->bprows, (size_t)c->g.vh * n) || dalloc(&c->slot_proj, n) || dalloc(&c->bpslow, 2)) {
        (void)hipGetLastError();
        return fail(TJ_ENOMEM, "cannot allocate %d projection slots of %dx%d", n, c->g.pw, c->g.ph);
    }
    c->cap = n;
    return ensure_iota(c, n);
}

static int ensure_weight_buffers(tj_ctx* c, int n)
{
    size_t need = tjk::weighting_workspace(c->g.pw, c->g.ph, n);
    if (need > c->wspace_size) {
        if (c->wspace) (void)hipFree(c->wspace);
        c->wspace = nullptr;
        c->wspace_size = 0;
        if (hipMalloc(&c->wspace, need) != hipSuccess) {
            (void)hipGetLastError();
            return fail(TJ_ENOMEM, "cannot allocate the weighting workspace (%zu bytes)", need);
        }
        c->wspace_size = need;
    }
    return TJ_OK;
}

static int check_range(const tj_ctx* c, int start_y, int end_y)
{
    if (start_y < 0 || end_y > c->g.vh || start_y > end_y)
        return fail(TJ_EINVAL, "row range [%d,%d) outside the volume height %d", start_y, end_y, c->g.vh);
    return TJ_OK;
}

// First detector row VoxelProjector3D.project computes (:338, :347).
int tj_fp_start_row(const tj_geometry& g, int start_y)
{
    const int offset = (g.proj_h > g.vol_h && g.scale == 1) ? (g.proj_h - g.vol_h) / 2 : 0;
    return (int)std::floor((start_y + offset) * g.scale + 0.5);
}

static int fp_start_row(const tj_ctx* c, int start_y) { return tj_fp_start_row(c->geo, start_y); }

// TjGeomDev from the ABI geometry: borders W + ACCURACY and H + ACCURACY are
// evaluated in float as in VoxelProjector3D.java:700-701.
void tj_geom_to_dev(const tj_geometry* geo, TjGeomDev* out)
{
    TjGeomDev& g = *out;
    g = TjGeomDev{};
    g.vw = geo->vol_w;
    g.vh = geo->vol_h;
    g.vd = geo->vol_d;
    g.pw = geo->proj_w;
    g.ph = geo->proj_h;
    g.positivity = geo->positivity ? 1 : 0;
    g.pc0 = geo->proj_center[0];
    g.pc1 = geo->proj_center[1];
    g.pc2 = geo->proj_center[2];
    g.pc3 = geo->proj_center[3];
    g.vcx = geo->vol_center[0];
    g.vcy = geo->vol_center[1];
    g.vcz = geo->vol_center[2];
    g.scale = geo->scale;
    g.bxmax = (double)((float)g.vw + 1e-10f);
    g.bymax = (double)((float)g.vh + 1e-10f);
}

static hipEvent_t ev(tj_ctx* c, size_t k)
{
    while (c->events.size() <= k) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->events.push_back(e);
    }
    return c->events[k];
}

// Forward projection of `nslot` slots whose projection indices are in device
// array `slots` (host copy `hslots`), into sproj/snorm.
static int do_project(tj_ctx* c, const int* slots, const std::vector<int>& hslots, int start_y)
{
    const int nslot = (int)hslots.size();
    const int start = fp_start_row(c, start_y);
    const int kernel = pick_fp_kernel(c, hslots, start);
    TJ_HIP(tjk::fp(c->stream, c->g, c->vol, c->pp, slots, nslot, start, kernel, c->pflag, c->norm_tab, c->sproj,
                   c->snorm, c->tune.fp_rows, c->tune.fp_rays, c->tune.fp_zchunk, c->tune.fp_groups,
                   c->tune.fp_prefetch == 0));
    if (c->geo.long_object_comp) TJ_HIP(tjk::norm_max_fill(c->stream, (int)c->npx, nslot, c->snorm));
    return TJ_OK;
}

static int do_residual(tj_ctx* c, const int* slots, int nslot, double factor, double* err_out)
{
    TJ_HIP(tjk::residual(c->stream, c->g.pw, c->g.ph, nslot, c->exps, slots, c->sproj, c->snorm, factor, c->sdiff,
                         c->row_err, c->rowflag, err_out));
    return TJ_OK;
}

// The volume the backprojection adds into: the reconstruction, or the
// Bayesian projector's distance volume (BayesianVoxelProjector3D.backproject
// :75-110), whose finiteness is not tracked (the checked update path).
static float* bp_target(tj_ctx* c) { return c->bayes ? c->dist : c->vol; }
static int bp_finite(tj_ctx* c) { return (c->bayes || c->any_nonfinite) ? 0 : 1; }

static int do_backproject(tj_ctx* c, const int* slots, const std::vector<int>& hslots, const float* images,
                          const int* image_index, int start_y, int end_y)
{
    const int kernel = pick_bp_kernel(c, hslots);
    if (kernel == 3)
        if (int rc = ensure_bpairs(c, (int)hslots.size())) return rc;
    TJ_HIP(tjk::bp(c->stream, c->g, bp_target(c), c->pp, slots, (int)hslots.size(), images, image_index,
                   c->rowflag, c->bprows, start_y, end_y, kernel, c->tune.bp_rows, c->tune.bp_zchunk,
                   c->tune.bp_group, bp_finite(c), c->bpairs, nullptr, c->bpslow, c->tune.bp_fast));
    return TJ_OK;
}

// The error-volume half of VoxelProjector3D.backproject (:181-195): the
// differences of the `nslot` slots squared in place (convertDiffToErrorImage,
// :315-324) and backprojected into the error volume, image by image in slot
// order, over rows [start_y, end_y).  TomoJ squares each difference image on
// the calling thread while the reconstruction's backprojection tasks of the
// same image may still be reading it; here the reconstruction's
// backprojection of the subset has been issued first on the same stream, so
// it always reads the differences themselves.  TomoJ's error tasks step
// `yy + increment` without the endY clamp the reconstruction tasks use; rows
// past endY are not touched here.
static int do_error_backproject(tj_ctx* c, const int* slots, const std::vector<int>& hslots, int start_y,
                                int end_y)
{
    const int n = (int)hslots.size();
    if (!c->errvol || n == 0) return TJ_OK;
    TJ_HIP(tjk::square_diffs(c->stream, c->sdiff, c->npx * (size_t)n));
    TJ_HIP(tjk::row_flags(c->stream, c->g.pw, c->g.ph, n, c->sdiff, c->iota, c->rowflag));
    const int kernel = pick_bp_kernel(c, hslots);
    if (kernel == 3)
        if (int rc = ensure_bpairs(c, n)) return rc;
    // the error volume holds sums of squares with the NaN/Inf scrub after
    // every image, so it is always finite
    TJ_HIP(tjk::bp(c->stream, c->g, c->errvol, c->pp, slots, n, c->sdiff, c->iota, c->rowflag, c->bprows, start_y,
                   end_y, kernel, c->tune.bp_rows, c->tune.bp_zchunk, c->tune.bp_group, 1, c->bpairs, nullptr,
                   c->bpslow, c->tune.bp_fast));
    return TJ_OK;
}

// TomoReconstruction2.orderShuffleART (:735-781) with the random first pick
// replaced by `first`: each next image is the remaining one whose row 2 of the
// Euler matrix (euler, n x 9) has the smallest |dot product| with the previous
// one's.
extern "C" void tj_order_shuffle(const double* euler, int n, int first, int* order)
{
    std::vector<char> todo(n, 1);
    int cc = 0, im = 0;
    while (cc < n) {
        if (cc == 0) {
            im = first;
        } else {
            double mini = DBL_MAX;
            int omini = -1;
            const double* a = euler + 9 * (size_t)im + 6;
            for (int o = 0; o < n; ++o) {
                if (!todo[o]) continue;
                const double* b = euler + 9 * (size_t)o + 6;
                double sum = 0;
                sum += a[0] * b[0];
                sum += a[1] * b[1];
                sum += a[2] * b[2];
                const double sp = std::fabs(sum);
                if (sp < mini) {
                    mini = sp;
                    omini = o;
                }
            }
            im = omini;
        }
        if (todo[im]) {
            order[cc] = im;
            todo[im] = 0;
            ++cc;
        }
    }
}

// The images a fused loop walks: the available list, or all.
static int num_available(const tj_ctx* c) { return c->avail.empty() ? c->nproj : (int)c->avail.size(); }

static int check_available(tj_ctx* c, const char* who)
{
    if ((int)c->avail.size() > c->nproj)
        return fail(TJ_EINVAL, "%s: %zu available images of %d", who, c->avail.size(), c->nproj);
    for (int v : c->avail)
        if (v < 0 || v >= c->nproj) return fail(TJ_EINVAL, "%s: available image %d of %d", who, v, c->nproj);
    return TJ_OK;
}

// orderShuffleART (:735-781) over the available images: the chain runs over
// positions 0..nav-1 with the Euler matrix of image im (getEulerMatrix(im),
// :741; not of the image the position names), and emits availableImages[im].
static void order_shuffle(const tj_ctx* c, int first, std::vector<int>& order)
{
    const int n = num_available(c);
    std::vector<double> e((size_t)9 * n);
    for (int p = 0; p < n; ++p)
        for (int k = 0; k < 9; ++k) e[(size_t)9 * p + k] = c->hpp[p].e[k];
    order.assign(n, 0);
    if (n > 0) tj_order_shuffle(e.data(), n, first, order.data());
    if (!c->avail.empty())
        for (int& o : order) o = c->avail[o];
}

extern "C" {

const char* tj_last_error(void) { return g_last_error.c_str(); }
const char* tj_version(void) { return "tomoj-mi355x 0.1 (gfx950)"; }

int tj_copy_probe(int device, size_t bytes, int repeats, double* ms)
{
    if (bytes < 16 || repeats < 1) return fail(TJ_EINVAL, "tj_copy_probe: bad size");
    TJ_HIP(hipSetDevice(device));
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < repeats && e == hipSuccess; ++r) e = tjk::copy(st, a, b, bytes);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / repeats;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_copy_probe: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_device_info(int device, char* buf, int len)
{
    if (!buf || len <= 0) return fail(TJ_EINVAL, "tj_device_info: no buffer");
    hipDeviceProp_t prop;
    TJ_HIP(hipGetDeviceProperties(&prop, device));
    int lds_optin = 0;
    (void)hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device);
    snprintf(buf, len,
             "{\"name\": \"%s\", \"arch\": \"%s\", \"cus\": %d, \"clock_mhz\": %d, "
             "\"lds_per_block\": %zu, \"lds_per_block_optin\": %d, \"lds_per_cu\": %zu, "
             "\"hbm_bytes\": %zu, \"l2_bytes\": %d}",
             prop.name, prop.gcnArchName, prop.multiProcessorCount, prop.clockRate / 1000, prop.sharedMemPerBlock,
             lds_optin, prop.maxSharedMemoryPerMultiProcessor, prop.totalGlobalMem, prop.l2CacheSize);
    return TJ_OK;
}

int tj_max_volume_bytes(int device, size_t* bytes)
{
    if (!bytes) return fail(TJ_EINVAL, "tj_max_volume_bytes: null argument");
    hipDeviceProp_t prop;
    TJ_HIP(hipGetDeviceProperties(&prop, device));
    // min(0.4 x global memory, largest allocation); HIP allocates up to the
    // whole memory, so the first term decides
    *bytes = (size_t)(prop.totalGlobalMem * 0.4);
    return TJ_OK;
}

void tj_geometry_default(tj_geometry* g, int vw, int vh, int vd, int pw, int ph)
{
    std::memset(g, 0, sizeof *g);
    g->vol_w = vw;
    g->vol_h = vh;
    g->vol_d = vd;
    g->proj_w = pw;
    g->proj_h = ph;
    g->proj_center[0] = (pw - 1.0) / 2;
    g->proj_center[1] = (ph - 1.0) / 2;
    g->proj_center[2] = (pw - 1) / 2.0;
    g->proj_center[3] = (ph - 1) / 2.0;
    g->vol_center[0] = (vw - 1) / 2.0;
    g->vol_center[1] = (vh - 1) / 2.0;
    g->vol_center[2] = (vd - 1) / 2.0;
    g->scale = 1.0;
    g->storage = TJ_STORAGE_F32;
}

int tj_open(int device, const tj_geometry* geo, tj_ctx** out)
{
    if (!geo || !out) return fail(TJ_EINVAL, "tj_open: null argument");
    *out = nullptr;
    if (geo->vol_w <= 0 || geo->vol_h <= 0 || geo->vol_d <= 0 || geo->proj_w <= 0 || geo->proj_h <= 0)
        return fail(TJ_EINVAL, "tj_open: non-positive size");
    if (!(geo->scale > 0)) return fail(TJ_EINVAL, "tj_open: scale must be > 0");
    if (geo->storage != TJ_STORAGE_F32) return fail(TJ_EUNSUP, "tj_open: storage type %d not supported", geo->storage);
    int ndev = 0;
    TJ_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(TJ_EINVAL, "tj_open: device %d of %d", device, ndev);
    TJ_HIP(hipSetDevice(device));
    int max_lds = 65536;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess)
        tjk::set_max_lds(max_lds);
    tj_ctx* c = new tj_ctx();
    c->device = device;
    c->geo = *geo;
    tj_geom_to_dev(geo, &c->g);
    const TjGeomDev& g = c->g;
    c->npx = (size_t)g.pw * g.ph;
    c->plane = (size_t)g.vd * g.vw;
    c->nvox = c->plane * g.vh;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        dalloc(&c->vol, c->nvox) != hipSuccess || dalloc(&c->pflag, (size_t)g.vh + 1) != hipSuccess) {
        (void)hipGetLastError();
        ctx_free(c);
        delete c;
        return fail(TJ_ENOMEM, "tj_open: cannot allocate a %dx%dx%d volume", g.vw, g.vh, g.vd);
    }
    if (hipMemsetAsync(c->vol, 0, c->nvox * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->pflag, 0, g.vh + 1, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        ctx_free(c);
        delete c;
        return fail(TJ_EDEVICE, "tj_open: clearing the volume failed");
    }
    *out = c;
    return TJ_OK;
}

int tj_close(tj_ctx* c)
{
    if (!c) return TJ_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ctx_free(c);
    delete c;
    return TJ_OK;
}

int tj_set_eulers(tj_ctx* c, const double* euler, int n)
{
    if (!c || !euler) return fail(TJ_EINVAL, "tj_set_eulers: null argument");
    if (n != c->nproj) return fail(TJ_EINVAL, "tj_set_eulers: %d matrices for %d projections", n, c->nproj);
    TJ_HIP(hipSetDevice(c->device));
    c->hpp.resize(n);
    for (int k = 0; k < n; ++k) c->hpp[k] = tj_make_proj(euler + 9 * k);
    TJ_HIP(hipMemcpy(c->pp, c->hpp.data(), sizeof(TjProj) * n, hipMemcpyHostToDevice));
    TJ_HIP(tjk::norm_table(c->stream, c->g, c->pp, n, c->norm_tab));
    TJ_HIP(hipStreamSynchronize(c->stream));
    return TJ_OK;
}

static int alloc_projections(tj_ctx* c, int n, const double* euler)
{
    dfree(c->exps);
    dfree(c->pp);
    dfree(c->order);
    dfree(c->norm_tab);
    c->nproj = 0;
    c->added.clear();
    c->have_diff = c->have_proj = false;
    if (dalloc(&c->exps, c->npx * n) || dalloc(&c->pp, n) || dalloc(&c->order, n) ||
        dalloc(&c->norm_tab, (size_t)c->g.pw * n)) {
        (void)hipGetLastError();
        return fail(TJ_ENOMEM, "cannot allocate %d projections of %dx%d", n, c->g.pw, c->g.ph);
    }
    c->nproj = n;
    int rc = tj_set_eulers(c, euler, n);
    if (rc) return rc;
    return ensure_iota(c, n);
}

int tj_load_projections(tj_ctx* c, const float* stack, int n, const double* euler)
{
    if (!c || !stack || !euler || n <= 0) return fail(TJ_EINVAL, "tj_load_projections: bad argument");
    TJ_HIP(hipSetDevice(c->device));
    int rc = alloc_projections(c, n, euler);
    if (rc) return rc;
    TJ_HIP(hipMemcpy(c->exps, stack, sizeof(float) * c->npx * n, hipMemcpyHostToDevice));
    return TJ_OK;
}

// TiltSeries.getImageStatistics ROI columns (TiltSeries.java:813-828) with the
// identity alignment: Math.round(W cos(tilt)) (at least min_size) columns,
// offset by integer division; the double round trip (x - cx) + cx of the Java
// code is kept.
static void roi_columns(int width, double tilt_deg, int min_size, int* x0, int* x1)
{
    const double rad = tilt_deg / 180.0 * M_PI;
    const long r = (long)std::floor(width * std::cos(rad) + 0.5);
    const int roisx = (int)(r > min_size ? r : min_size);
    const double cx = (width - 1.0) / 2;
    const double p0 = (double)((width - roisx) / 2) - cx;
    *x0 = (int)(p0 + cx);
    *x1 = (int)((p0 + roisx) + cx);
}

int tj_normalize_projections(tj_ctx* c, const double* tilt_deg, int min_size, double* stats)
{
    if (!c || !tilt_deg) return fail(TJ_EINVAL, "tj_normalize_projections: null argument");
    if (c->nproj <= 0) return fail(TJ_ESTATE, "tj_normalize_projections: no resident projections");
    TJ_HIP(hipSetDevice(c->device));
    std::vector<int> roi(4 * (size_t)c->nproj);
    for (int p = 0; p < c->nproj; ++p) {
        roi_columns(c->g.pw, tilt_deg[p], min_size, &roi[4 * p], &roi[4 * p + 1]);
        roi[4 * p + 2] = 0;
        roi[4 * p + 3] = c->g.ph;
    }
    int* droi = nullptr;
    double* dstats = nullptr;
    hipError_t e = dalloc(&droi, roi.size());
    if (e == hipSuccess) e = dalloc(&dstats, 2 * (size_t)c->nproj);
    if (e == hipSuccess) e = hipMemcpyAsync(droi, roi.data(), sizeof(int) * roi.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = tjk::normalize(c->stream, c->exps, c->g.pw, c->g.ph, c->nproj, droi, dstats);
    if (e == hipSuccess && stats)
        e = hipMemcpyAsync(stats, dstats, sizeof(double) * 2 * c->nproj, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(droi);
    dfree(dstats);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_normalize_projections: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_threshold_projections(tj_ctx* c, double threshold, double hysteresis, int fill)
{
    if (!c) return fail(TJ_EINVAL, "tj_threshold_projections: null context");
    if (c->nproj <= 0) return fail(TJ_ESTATE, "tj_threshold_projections: no resident projections");
    if (fill < TJ_FILL_NONE || fill > TJ_FILL_NAN) return fail(TJ_EINVAL, "tj_threshold_projections: fill %d", fill);
    TJ_HIP(hipSetDevice(c->device));
    double* lv = nullptr;
    uint8_t* q = nullptr;
    int* changed = nullptr;
    hipError_t e = dalloc(&lv, 4 * (size_t)c->nproj);
    if (e == hipSuccess) e = dalloc(&q, c->npx * c->nproj);
    if (e == hipSuccess) e = dalloc(&changed, 1);
    if (e == hipSuccess)
        e = tjk::threshold(c->stream, c->exps, c->g.pw, c->g.ph, c->nproj, threshold, hysteresis, fill, lv, q,
                           changed);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(lv);
    dfree(q);
    dfree(changed);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? TJ_ENOMEM : TJ_EDEVICE, "tj_threshold_projections: %s",
                    hipGetErrorString(e));
    }
    return TJ_OK;
}

// java.awt.geom.AffineTransform.createInverse by the transform's state (the
// same partition of the matrix as AffineTransform.updateState): translation,
// scale and pure-shear states invert entry by entry, the general state by the
// determinant.  m: m00, m10, m01, m11, m02, m12 (AffineTransform.getMatrix
// order).  Returns false for a non-invertible transform.
static bool affine_inverse(const double* m, double* r)
{
    const double m00 = m[0], m10 = m[1], m01 = m[2], m11 = m[3], m02 = m[4], m12 = m[5];
    const bool trans = !(m02 == 0.0 && m12 == 0.0);
    auto set = [&](double a, double b, double c2, double d, double e, double f) {
        r[0] = a; r[1] = b; r[2] = c2; r[3] = d; r[4] = e; r[5] = f;
    };
    if (m01 == 0.0 && m10 == 0.0) {
        if (m00 == 1.0 && m11 == 1.0) {
            if (trans) set(1.0, 0.0, 0.0, 1.0, -m02, -m12);
            else set(1.0, 0.0, 0.0, 1.0, 0.0, 0.0);
            return true;
        }
        if (m00 == 0.0 || m11 == 0.0) return false;
        if (trans) set(1.0 / m00, 0.0, 0.0, 1.0 / m11, -m02 / m00, -m12 / m11);
        else set(1.0 / m00, 0.0, 0.0, 1.0 / m11, 0.0, 0.0);
        return true;
    }
    if (m00 == 0.0 && m11 == 0.0) {
        if (m01 == 0.0 || m10 == 0.0) return false;
        if (trans) set(0.0, 1.0 / m01, 1.0 / m10, 0.0, -m12 / m10, -m02 / m01);
        else set(0.0, 1.0 / m01, 1.0 / m10, 0.0, 0.0, 0.0);
        return true;
    }
    const double det = m00 * m11 - m01 * m10;
    if (std::fabs(det) <= 4.9406564584124654e-324) return false;  // Double.MIN_VALUE
    if (trans)
        set(m11 / det, -m10 / det, -m01 / det, m00 / det, (m01 * m12 - m11 * m02) / det,
            (m10 * m02 - m00 * m12) / det);
    else
        set(m11 / det, -m10 / det, -m01 / det, m00 / det, 0.0, 0.0);
    return true;
}

// TiltSeries.getImageStatistics' polygon (TiltSeries.java:809-835): the ROI
// rectangle of the identity case, its corners mapped by T and cast to int.  For
// a transform without rotation or shear the polygon is the axis-aligned
// rectangle [min x, max x) x [min y, max y).
static void roi_poly_affine(int width, int height, double tilt_deg, int min_size, const double* m, int* xs,
                            int* ys);

static void roi_rect_affine(int width, int height, double tilt_deg, int min_size, const double* m, int* r4)
{
    int xs[4], ys[4];
    roi_poly_affine(width, height, tilt_deg, min_size, m, xs, ys);
    r4[0] = std::min(std::min(xs[0], xs[1]), std::min(xs[2], xs[3]));
    r4[1] = std::max(std::max(xs[0], xs[1]), std::max(xs[2], xs[3]));
    r4[2] = std::min(std::min(ys[0], ys[1]), std::min(ys[2], ys[3]));
    r4[3] = std::max(std::max(ys[0], ys[1]), std::max(ys[2], ys[3]));
}

// The polygon's int corners.
static void roi_poly_affine(int width, int height, double tilt_deg, int min_size, const double* m, int* xs,
                            int* ys)
{
    const double rad = tilt_deg / 180.0 * M_PI;
    const long rr = (long)std::floor(width * std::cos(rad) + 0.5);
    const int roisx = (int)(rr > min_size ? rr : min_size);
    const double cx = (width - 1.0) / 2, cy = (height - 1.0) / 2;
    double pts[8];
    pts[0] = (double)((width - roisx) / 2) - cx;  // integer division, then - cx
    pts[1] = -cy;
    pts[2] = pts[0] + roisx;
    pts[3] = pts[1];
    pts[4] = pts[2];
    pts[5] = height - cy;
    pts[6] = pts[0];
    pts[7] = pts[5];
    for (int k = 0; k < 4; ++k) {
        const double x = pts[2 * k], y = pts[2 * k + 1];
        xs[k] = (int)((m[0] * x + m[2] * y + m[4]) + cx);
        ys[k] = (int)((m[1] * x + m[3] * y + m[5]) + cy);
    }
}

int tj_align_projections(tj_ctx* c, const double* transforms, const double* tilt_deg, int min_size, int normalize,
                         int fill, const double* stats_in, double* stats_out)
{
    if (!c || !transforms) return fail(TJ_EINVAL, "tj_align_projections: null argument");
    if (c->nproj <= 0) return fail(TJ_ESTATE, "tj_align_projections: no resident projections");
    if (fill < TJ_FILL_NONE || fill > TJ_FILL_NAN) return fail(TJ_EINVAL, "tj_align_projections: fill %d", fill);
    const int n = c->nproj, w = c->g.pw, h = c->g.ph;
    const bool need_stats = normalize || fill == TJ_FILL_AVG;
    std::vector<double> inv(6 * (size_t)n);
    // ROI of each image's statistics: a rectangle [x0,x1) x [y0,y1) when the
    // transform neither rotates nor shears (roi), else the polygon's corners
    // (poly); `rotated` says which.
    std::vector<int> ident(n), roi(4 * (size_t)n), poly(8 * (size_t)n);
    bool rotated = false;
    for (int p = 0; p < n; ++p) {
        const double* m = transforms + 6 * (size_t)p;
        ident[p] = (m[0] == 1.0 && m[1] == 0.0 && m[2] == 0.0 && m[3] == 1.0 && m[4] == 0.0 && m[5] == 0.0);
        if (!affine_inverse(m, &inv[6 * (size_t)p]))
            return fail(TJ_EINVAL, "tj_align_projections: transform %d is not invertible", p);
        if (need_stats && !stats_in) {
            if (!tilt_deg) return fail(TJ_EINVAL, "tj_align_projections: tilt angles needed for the statistics");
            roi_rect_affine(w, h, tilt_deg[p], min_size, m, &roi[4 * (size_t)p]);
            roi_poly_affine(w, h, tilt_deg[p], min_size, m, &poly[8 * (size_t)p], &poly[8 * (size_t)p + 4]);
            rotated |= (m[1] != 0.0 || m[2] != 0.0);
        }
    }
    TJ_HIP(hipSetDevice(c->device));
    double *dinv = nullptr, *dstats = nullptr;
    int *dident = nullptr, *droi = nullptr;
    float* orig = nullptr;
    hipError_t e = dalloc(&dinv, inv.size());
    if (e == hipSuccess) e = dalloc(&dident, ident.size());
    if (e == hipSuccess) e = dalloc(&dstats, 2 * (size_t)n);
    if (e == hipSuccess) e = dalloc(&orig, c->npx * n);
    if (e == hipSuccess) e = hipMemcpyAsync(dinv, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dident, ident.data(), sizeof(int) * ident.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(orig, c->exps, sizeof(float) * c->npx * n, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && need_stats) {
        if (stats_in) {
            e = hipMemcpyAsync(dstats, stats_in, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream);
        } else {
            // the polygon test gives the rectangle's pixels for rotation-free
            // transforms; those keep the rectangle kernel
            const std::vector<int>& r = rotated ? poly : roi;
            e = dalloc(&droi, r.size());
            if (e == hipSuccess)
                e = hipMemcpyAsync(droi, r.data(), sizeof(int) * r.size(), hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess)
                e = rotated ? tjk::poly_stats(c->stream, orig, w, h, n, droi, dstats)
                            : tjk::roi_stats(c->stream, orig, w, h, n, droi, dstats);
        }
    }
    if (e == hipSuccess)
        e = tjk::affine(c->stream, orig, c->exps, w, h, n, dinv, dident, need_stats ? dstats : nullptr,
                        normalize ? 1 : 0, fill);
    if (e == hipSuccess && stats_out && need_stats)
        e = hipMemcpyAsync(stats_out, dstats, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(dinv);
    dfree(dident);
    dfree(dstats);
    dfree(droi);
    dfree(orig);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_align_projections: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_get_loaded_projection(tj_ctx* c, int index, float* out)
{
    if (!c || !out) return fail(TJ_EINVAL, "tj_get_loaded_projection: null argument");
    if (index < 0 || index >= c->nproj) return fail(TJ_EINVAL, "projection index %d of %d", index, c->nproj);
    TJ_HIP(hipSetDevice(c->device));
    TJ_HIP(hipMemcpy(out, c->exps + (size_t)index * c->npx, sizeof(float) * c->npx, hipMemcpyDeviceToHost));
    return TJ_OK;
}

int tj_set_volume(tj_ctx* c, const float* zyx, int start_y, int end_y)
{
    if (!c || !zyx) return fail(TJ_EINVAL, "tj_set_volume: null argument");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    const int nr = end_y - start_y;
    if (nr == 0) return TJ_OK;
    TJ_HIP(hipSetDevice(c->device));
    const TjGeomDev& g = c->g;
    float* stage = nullptr;
    TJ_HIP(dalloc(&stage, (size_t)g.vd * nr * g.vw));
    hipError_t e = hipMemcpy2DAsync(stage, sizeof(float) * nr * g.vw, zyx + (size_t)start_y * g.vw,
                                    sizeof(float) * g.vh * g.vw, sizeof(float) * nr * g.vw, g.vd,
                                    hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = tjk::zyx_to_yzx(c->stream, stage, c->vol, g.vw, g.vd, nr, start_y);
    if (e == hipSuccess) e = tjk::plane_flags(c->stream, c->vol, c->plane, start_y, nr, c->pflag);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_set_volume: %s", hipGetErrorString(e));
    std::vector<uint8_t> hf(g.vh + 1);
    TJ_HIP(hipMemcpy(hf.data(), c->pflag, g.vh + 1, hipMemcpyDeviceToHost));
    c->any_nonfinite = false;
    for (uint8_t f : hf) c->any_nonfinite |= (f != 0);
    return TJ_OK;
}

int tj_get_volume(tj_ctx* c, float* zyx, int start_y, int end_y)
{
    if (!c || !zyx) return fail(TJ_EINVAL, "tj_get_volume: null argument");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    const int nr = end_y - start_y;
    if (nr == 0) return TJ_OK;
    TJ_HIP(hipSetDevice(c->device));
    const TjGeomDev& g = c->g;
    float* stage = nullptr;
    TJ_HIP(dalloc(&stage, (size_t)g.vd * nr * g.vw));
    hipError_t e = tjk::yzx_to_zyx(c->stream, c->vol, stage, g.vw, g.vd, nr, start_y);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(zyx + (size_t)start_y * g.vw, sizeof(float) * g.vh * g.vw, stage,
                             sizeof(float) * nr * g.vw, sizeof(float) * nr * g.vw, g.vd, hipMemcpyDeviceToHost,
                             c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_get_volume: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_reset_volume(tj_ctx* c)
{
    if (!c) return fail(TJ_EINVAL, "tj_reset_volume: null context");
    TJ_HIP(hipSetDevice(c->device));
    TJ_HIP(hipMemsetAsync(c->vol, 0, c->nvox * sizeof(float), c->stream));
    TJ_HIP(hipMemsetAsync(c->pflag, 0, c->g.vh + 1, c->stream));
    TJ_HIP(hipStreamSynchronize(c->stream));
    c->any_nonfinite = false;
    return TJ_OK;
}

int tj_set_positivity(tj_ctx* c, int on)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_positivity: null context");
    c->geo.positivity = on ? 1 : 0;
    c->g.positivity = on ? 1 : 0;
    return TJ_OK;
}

int tj_set_error_volume(tj_ctx* c, int on)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_error_volume: null context");
    TJ_HIP(hipSetDevice(c->device));
    if (!on) {
        TJ_HIP(hipStreamSynchronize(c->stream));
        dfree(c->errvol);
        return TJ_OK;
    }
    if (!c->errvol) {
        if (dalloc(&c->errvol, c->nvox) != hipSuccess) {
            (void)hipGetLastError();
            c->errvol = nullptr;
            return fail(TJ_ENOMEM, "tj_set_error_volume: cannot allocate a %dx%dx%d error volume", c->g.vw, c->g.vh,
                        c->g.vd);
        }
    }
    return tj_clear_error_volume(c);
}

int tj_clear_error_volume(tj_ctx* c)
{
    if (!c) return fail(TJ_EINVAL, "tj_clear_error_volume: null context");
    if (!c->errvol) return TJ_OK;
    TJ_HIP(hipSetDevice(c->device));
    TJ_HIP(hipMemsetAsync(c->errvol, 0, c->nvox * sizeof(float), c->stream));
    TJ_HIP(hipStreamSynchronize(c->stream));
    return TJ_OK;
}

int tj_get_error_volume(tj_ctx* c, float* zyx)
{
    if (!c || !zyx) return fail(TJ_EINVAL, "tj_get_error_volume: null argument");
    if (!c->errvol) return fail(TJ_ESTATE, "tj_get_error_volume: no error volume (tj_set_error_volume)");
    TJ_HIP(hipSetDevice(c->device));
    const TjGeomDev& g = c->g;
    float* stage = nullptr;
    TJ_HIP(dalloc(&stage, c->nvox));
    hipError_t e = tjk::yzx_to_zyx(c->stream, c->errvol, stage, g.vw, g.vd, g.vh, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(zyx, stage, c->nvox * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_get_error_volume: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_add_projection(tj_ctx* c, int index)
{
    if (!c) return fail(TJ_EINVAL, "tj_add_projection: null context");
    if (index < 0 || index >= c->nproj)
        return fail(TJ_EINVAL, "tj_add_projection: index %d outside the %d resident projections", index, c->nproj);
    c->added.push_back(index);
    return TJ_OK;
}

int tj_clear_projections(tj_ctx* c)
{
    if (!c) return fail(TJ_EINVAL, "tj_clear_projections: null context");
    c->added.clear();
    c->have_diff = c->have_proj = false;
    return TJ_OK;
}

int tj_num_added(tj_ctx* c) { return c ? (int)c->added.size() : 0; }

// The added indices on the device (slot_proj); a list already there (the
// backprojection after projectAndDifference of the same subset) is not sent
// again.
static int upload_added(tj_ctx* c)
{
    const int cap = c->cap;
    int rc = ensure_slots(c, (int)c->added.size());
    if (rc) return rc;
    if (c->cap == cap && c->uploaded == c->added) return TJ_OK;
    TJ_HIP(hipMemcpyAsync(c->slot_proj, c->added.data(), sizeof(int) * c->added.size(), hipMemcpyHostToDevice,
                          c->stream));
    c->uploaded = c->added;
    return TJ_OK;
}

int tj_project(tj_ctx* c, int start_y, int end_y)
{
    if (!c) return fail(TJ_EINVAL, "tj_project: null context");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    TJ_HIP(hipSetDevice(c->device));
    c->completion.store(0.0);
    if (c->added.empty()) {
        c->have_proj = true;
        return TJ_OK;
    }
    if ((rc = upload_added(c))) return rc;
    if ((rc = do_project(c, c->slot_proj, c->added, start_y))) return rc;
    TJ_HIP(hipStreamSynchronize(c->stream));
    c->have_proj = true;
    return TJ_OK;
}

int tj_project_and_difference(tj_ctx* c, double factor, int start_y, int end_y, double* errors)
{
    if (!c) return fail(TJ_EINVAL, "tj_project_and_difference: null context");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    TJ_HIP(hipSetDevice(c->device));
    c->completion.store(0.0);
    const int n = (int)c->added.size();
    if (n == 0) {
        c->have_proj = c->have_diff = true;
        return TJ_OK;
    }
    if ((rc = upload_added(c))) return rc;
    const int start = fp_start_row(c, start_y);
    if (pick_fp_kernel(c, c->added, start) == 3 && !c->geo.long_object_comp) {
        // the LDS forward kernel fused with the residual, writing the
        // projections and norms as well (getProjection after the call)
        TJ_HIP(tjk::fp_residual(c->stream, c->g, c->vol, c->pp, c->slot_proj, n, start, c->norm_tab,
                                c->tune.fp_rays, c->tune.fp_zchunk, c->exps, factor, c->sdiff, c->row_err, c->rbad,
                                c->rowflag, c->err, nullptr, c->tune.fp_groups, nullptr, c->tune.fp_prefetch == 0,
                                c->sproj, c->snorm));
    } else {
        if ((rc = do_project(c, c->slot_proj, c->added, start_y))) return rc;
        if ((rc = do_residual(c, c->slot_proj, n, factor, c->err))) return rc;
    }
    TJ_HIP(hipMemcpyAsync(c->h_err, c->err, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    TJ_HIP(hipStreamSynchronize(c->stream));
    if (errors) std::memcpy(errors, c->h_err, sizeof(double) * n);
    c->have_proj = c->have_diff = true;
    return TJ_OK;
}

int tj_last_error_sum(tj_ctx* c, double* som)
{
    if (!c || !som) return fail(TJ_EINVAL, "tj_last_error_sum: null argument");
    if (!c->have_diff) return fail(TJ_ESTATE, "no difference computed since the last clear");
    double s = 0;
    for (size_t k = 0; k < c->added.size(); ++k) s += c->h_err[k];
    *som = s;
    return TJ_OK;
}

static int get_slot_image(tj_ctx* c, int slot, const float* base, float* out, const char* what)
{
    if (slot < 0 || slot >= (int)c->added.size())
        return fail(TJ_EINVAL, "%s: slot %d of %zu added projections", what, slot, c->added.size());
    TJ_HIP(hipSetDevice(c->device));
    TJ_HIP(hipMemcpy(out, base + (size_t)slot * c->npx, sizeof(float) * c->npx, hipMemcpyDeviceToHost));
    return TJ_OK;
}

int tj_get_projection(tj_ctx* c, int slot, float* proj, float* norm)
{
    if (!c) return fail(TJ_EINVAL, "tj_get_projection: null context");
    if (!c->have_proj) return fail(TJ_ESTATE, "tj_get_projection: no projection computed since the last clear");
    int rc;
    if (proj && (rc = get_slot_image(c, slot, c->sproj, proj, "tj_get_projection"))) return rc;
    if (norm && (rc = get_slot_image(c, slot, c->snorm, norm, "tj_get_projection"))) return rc;
    return TJ_OK;
}

int tj_get_difference(tj_ctx* c, int slot, float* diff)
{
    if (!c || !diff) return fail(TJ_EINVAL, "tj_get_difference: null argument");
    if (!c->have_diff) return fail(TJ_ESTATE, "tj_get_difference: no difference computed since the last clear");
    return get_slot_image(c, slot, c->sdiff, diff, "tj_get_difference");
}

// Weighted (or plain) images of the added projections into swgt and their row
// flags; returns the image array the backprojection reads.
static int prepare_originals(tj_ctx* c, const int* slots, const std::vector<int>& hslots, const float** images,
                             const int** image_index)
{
    const int n = (int)hslots.size();
    if (c->weight_kind == 0) {
        *images = c->exps;
        *image_index = slots;
    } else {
        if (c->swgt_cap < n) {
            dfree(c->swgt);
            c->swgt_cap = 0;
            TJ_HIP(dalloc(&c->swgt, c->npx * n));
            c->swgt_cap = n;
        }
        int rc = ensure_weight_buffers(c, n);
        if (rc) return rc;
        TJ_HIP(tjk::weighting_apply(c->stream, c->g.pw, c->g.ph, n, c->exps, slots, c->mask, c->swgt, c->wspace,
                                    c->wspace_size));
        *images = c->swgt;
        *image_index = c->iota;
    }
    TJ_HIP(tjk::row_flags(c->stream, c->g.pw, c->g.ph, n, *images, *image_index, c->rowflag));
    return TJ_OK;
}

int tj_backproject(tj_ctx* c, int start_y, int end_y)
{
    if (!c) return fail(TJ_EINVAL, "tj_backproject: null context");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    TJ_HIP(hipSetDevice(c->device));
    c->completion.store(0.0);
    if (c->added.empty()) return TJ_OK;
    if ((rc = upload_added(c))) return rc;
    const float* images;
    const int* index;
    if (c->have_diff) {
        images = c->sdiff;
        index = c->iota;
    } else {
        // BayesianVoxelProjector3D.backproject only adds differences (:90)
        if (c->bayes) return TJ_OK;
        if ((rc = prepare_originals(c, c->slot_proj, c->added, &images, &index))) return rc;
    }
    if ((rc = do_backproject(c, c->slot_proj, c->added, images, index, start_y, end_y))) return rc;
    if (c->have_diff && (rc = do_error_backproject(c, c->slot_proj, c->added, start_y, end_y))) return rc;
    TJ_HIP(hipStreamSynchronize(c->stream));
    return TJ_OK;
}

int tj_set_weighting(tj_ctx* c, double diameter, const double* tilt_deg, const double* axis_deg, int n)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_weighting: null context");
    TJ_HIP(hipSetDevice(c->device));
    if (!(diameter > 0)) {
        c->weight_kind = 0;
        return TJ_OK;
    }
    if (!tilt_deg || n <= 0) return fail(TJ_EINVAL, "tj_set_weighting: no tilt angles");
    if (!c->mask) TJ_HIP(dalloc(&c->mask, c->npx));
    TJ_HIP(tjk::weighting_init(c->stream, c->g.pw, c->g.ph, diameter, tilt_deg, axis_deg, n, c->mask));
    c->weight_kind = 1;
    return TJ_OK;
}

int tj_set_ramp_filter(tj_ctx* c, int hamming, double cutoff)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_ramp_filter: null context");
    if (!(cutoff > 0 && cutoff <= 1)) return fail(TJ_EINVAL, "tj_set_ramp_filter: cutoff %g not in (0,1]", cutoff);
    TJ_HIP(hipSetDevice(c->device));
    if (!c->mask) TJ_HIP(dalloc(&c->mask, c->npx));
    if (c->nproj <= 0) return fail(TJ_ESTATE, "tj_set_ramp_filter: load the projections first");
    TJ_HIP(tjk::ramp_init(c->stream, c->g.pw, c->g.ph, hamming, cutoff, M_PI / c->nproj, c->mask));
    TJ_HIP(hipStreamSynchronize(c->stream));
    c->weight_kind = 2;
    return TJ_OK;
}

int tj_get_weighting_mask(tj_ctx* c, float* mask)
{
    if (!c || !mask) return fail(TJ_EINVAL, "tj_get_weighting_mask: null argument");
    if (c->weight_kind == 0) return fail(TJ_ESTATE, "tj_get_weighting_mask: no weighting filter set");
    TJ_HIP(hipSetDevice(c->device));
    TJ_HIP(hipMemcpy(mask, c->mask, sizeof(float) * c->npx, hipMemcpyDeviceToHost));
    return TJ_OK;
}

int tj_apply_weighting(tj_ctx* c, float* image)
{
    if (!c || !image) return fail(TJ_EINVAL, "tj_apply_weighting: null argument");
    if (c->weight_kind == 0) return fail(TJ_ESTATE, "tj_apply_weighting: no weighting filter set");
    TJ_HIP(hipSetDevice(c->device));
    int rc = ensure_slots(c, 1);
    if (rc) return rc;
    if ((rc = ensure_weight_buffers(c, 1))) return rc;
    float* tmp = nullptr;
    TJ_HIP(dalloc(&tmp, c->npx * 2));
    hipError_t e = hipMemcpy(tmp, image, sizeof(float) * c->npx, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = tjk::weighting_apply(c->stream, c->g.pw, c->g.ph, 1, tmp, c->iota, c->mask, tmp + c->npx, c->wspace,
                                 c->wspace_size);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(image, tmp + c->npx, sizeof(float) * c->npx, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_apply_weighting: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_set_tuning(tj_ctx* c, const tj_tuning* t)
{
    if (!c || !t) return fail(TJ_EINVAL, "tj_set_tuning: null argument");
    if (t->fp_kernel < 0 || t->fp_kernel > 3 || t->bp_kernel < 0 || t->bp_kernel > 3)
        return fail(TJ_EINVAL, "tj_set_tuning: kernel index out of range");
    if (t->fp_rows < 1 || t->bp_rows < 1 || t->fp_rays < 1 || t->bp_zchunk < 0 || t->fp_zchunk < 0 ||
        t->bp_group < 1)
        return fail(TJ_EINVAL, "tj_set_tuning: non-positive shape");
    if (t->bp_fast < 0 || t->bp_fast > 7) return fail(TJ_EINVAL, "tj_set_tuning: bp_fast out of range");
    if (t->fp_groups != 0 && t->fp_groups != 1 && t->fp_groups != 2 && t->fp_groups != 4)
        return fail(TJ_EINVAL, "tj_set_tuning: fp_groups must be 0, 1, 2 or 4");
    if (t->fp_prefetch != 0 && t->fp_prefetch != 1)
        return fail(TJ_EINVAL, "tj_set_tuning: fp_prefetch must be 0 or 1");
    c->tune = *t;
    return TJ_OK;
}

int tj_get_tuning(tj_ctx* c, tj_tuning* t)
{
    if (!c || !t) return fail(TJ_EINVAL, "tj_get_tuning: null argument");
    *t = c->tune;
    return TJ_OK;
}

// FistaProjector3D.startOfIteration (:62-67): keep the volume of the start of
// the iteration.
static int fista_start(tj_ctx* c)
{
    if (!c->fista) return TJ_OK;
    TJ_HIP(hipMemcpyAsync(c->prev, c->vol, c->nvox * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    return TJ_OK;
}

// Non-finite flags of every volume plane, after a whole-volume update that may
// overflow (FISTA extrapolation, TV denoising); `computed`: the update's own
// kernel has set them already.
static int refresh_plane_flags(tj_ctx* c, bool computed = false)
{
    if (!computed) TJ_HIP(tjk::plane_flags(c->stream, c->vol, c->plane, 0, c->g.vh, c->pflag));
    std::vector<uint8_t> hf(c->g.vh + 1);
    TJ_HIP(hipMemcpyAsync(hf.data(), c->pflag, c->g.vh + 1, hipMemcpyDeviceToHost, c->stream));
    TJ_HIP(hipStreamSynchronize(c->stream));
    c->any_nonfinite = false;
    for (uint8_t f : hf) c->any_nonfinite |= (f != 0);
    return TJ_OK;
}

// FistaProjector3D.endOfIteration (:43-59): t and the extrapolation factor on
// the host in double, the voxel update on the device.
static int fista_end(tj_ctx* c)
{
    if (!c->fista) return TJ_OK;
    const double t_1 = c->fista_t;
    c->fista_t = (1 + std::sqrt(1 + 4 * t_1 * t_1)) / 2;
    const double factor = (t_1 - 1) / c->fista_t;
    TJ_HIP(tjk::fista(c->stream, c->vol, c->prev, c->nvox, factor));
    return refresh_plane_flags(c);
}

static int ensure_tv(tj_ctx* c)
{
    if (c->tv_f && c->tv_p) return TJ_OK;
    if (!c->tv_f && dalloc(&c->tv_f, c->nvox) != hipSuccess) {
        (void)hipGetLastError();
        return fail(TJ_ENOMEM, "TV denoising: cannot allocate the input copy (%zu voxels)", c->nvox);
    }
    if (!c->tv_p && dalloc(&c->tv_p, 3 * c->nvox) != hipSuccess) {
        (void)hipGetLastError();
        dfree(c->tv_f);
        return fail(TJ_ENOMEM, "TV denoising: cannot allocate the dual field (%zu voxels x 3)", c->nvox);
    }
    return TJ_OK;
}

// TotalVariationDenoising_3D.denoise3D over the whole volume (every row, as
// TVMProjector3D passes rec's whole stack).
static int tv_run(tj_ctx* c, float theta, float g, float dt, int iterations)
{
    if (iterations <= 0) return TJ_OK;
    int rc = ensure_tv(c);
    if (rc) return rc;
    TJ_HIP(tjk::tv_denoise(c->stream, c->vol, c->tv_f, c->tv_p, c->g.vw, c->g.vh, c->g.vd, theta, g, dt,
                           iterations, c->pflag));
    return refresh_plane_flags(c, true);
}

// TVMProjector3D.endOfIteration (:135-157): dt is reset to the regularisation
// lambda before every denoising, so the tvmDt setting never reaches it.
static int tvm_end(tj_ctx* c)
{
    if (c->tvm_iter <= 0) return TJ_OK;
    return tv_run(c, c->tvm_theta, c->tvm_g, c->tvm_dt, c->tvm_iter);
}

static int ensure_dist(tj_ctx* c)
{
    if (c->dist) return TJ_OK;
    if (dalloc(&c->dist, c->nvox) != hipSuccess) {
        (void)hipGetLastError();
        return fail(TJ_ENOMEM, "Bayesian projector: cannot allocate the distance volume (%zu voxels)", c->nvox);
    }
    return TJ_OK;
}

// BayesianVoxelProjector3D.startOfIteration (:134-137): the distance volume
// restarts as -lambda times the convolved volume.
static int bayes_start(tj_ctx* c)
{
    if (!c->bayes) return TJ_OK;
    TJ_HIP(tjk::bayes_conv(c->stream, c->vol, c->dist, c->g.vw, c->g.vh, c->g.vd, -c->bayes_lambda));
    return TJ_OK;
}

// BayesianVoxelProjector3D.endOfIteration (:121-132).
static int bayes_end(tj_ctx* c)
{
    if (!c->bayes) return TJ_OK;
    TJ_HIP(tjk::bayes_end(c->stream, c->vol, c->dist, c->plane, c->g.vh, (float)c->bayes_alpha,
                          c->g.positivity, c->pflag));
    return refresh_plane_flags(c, true);
}

// BgARTVoxelProjector3D.collapseFilter (:30-78): the stack's 256-bin
// histogram on the device, the background level and thresholds on the host
// in Java's order, the collapse on the device.
static int bgart_collapse(tj_ctx* c)
{
    TJ_HIP(tjk::stack_stats(c->stream, c->vol, c->nvox, c->bg_mm, c->bg_hist));
    double mm[2];
    int hist[256];
    TJ_HIP(hipMemcpyAsync(mm, c->bg_mm, sizeof mm, hipMemcpyDeviceToHost, c->stream));
    TJ_HIP(hipMemcpyAsync(hist, c->bg_hist, sizeof hist, hipMemcpyDeviceToHost, c->stream));
    TJ_HIP(hipStreamSynchronize(c->stream));
    const double histmin = mm[0], histmax = mm[1];
    const double bin = (histmax - histmin) / 256;
    double best = 0;
    int bi = 0;
    for (int i = 0; i < 256; ++i)
        if (hist[i] > best) {
            best = hist[i];
            bi = i;
        }
    const double mv = std::isnan(c->bg_value) ? histmin + bi * bin : c->bg_value;
    double sigma = 0;
    int count = 0;
    for (int i = 0; i < bi; ++i) {
        sigma += (histmin + i * bin - mv) * (histmin + i * bin - mv) * hist[i];
        count += hist[i];
    }
    sigma = std::sqrt(sigma / count);
    TJ_HIP(tjk::collapse(c->stream, c->vol, c->nvox, mv - c->bg_k * sigma, mv + c->bg_k * sigma, (float)mv,
                         c->bg_dark));
    return refresh_plane_flags(c);
}

// BgARTVoxelProjector3D.startOfIteration (:123-151): the cross median.
static int bgart_start(tj_ctx* c)
{
    if (!c->bgart || !c->bg_median) return TJ_OK;
    TJ_HIP(hipMemcpyAsync(c->bg_tmp, c->vol, c->nvox * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    TJ_HIP(tjk::median_cross(c->stream, c->bg_tmp, c->vol, c->g.vw, c->g.vh, c->g.vd));
    return refresh_plane_flags(c);
}

// BgARTVoxelProjector3D.endOfIteration (:119-121): the collapse filter.
static int bgart_end(tj_ctx* c)
{
    if (!c->bgart) return TJ_OK;
    return bgart_collapse(c);
}

// FistaProjector3D wraps the working projector: its startOfIteration keeps the
// volume, then runs the wrapped one's (:62-67); its endOfIteration runs the
// wrapped one's first (:44).  TVM, Bayesian and BgART are never on together.
static int start_hooks(tj_ctx* c)
{
    int rc = fista_start(c);
    if (rc) return rc;
    if ((rc = bayes_start(c))) return rc;
    return bgart_start(c);
}

static int end_hooks(tj_ctx* c)
{
    int rc = tvm_end(c);
    if (rc) return rc;
    if ((rc = bayes_end(c))) return rc;
    if ((rc = bgart_end(c))) return rc;
    return fista_end(c);
}

int tj_set_bgart(tj_ctx* c, int on, double k, int dark_bg, int median, double bg_value)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_bgart: null context");
    if (on && (c->tvm_iter > 0 || c->bayes)) return fail(TJ_ESTATE, "tj_set_bgart: another regulariser is on");
    TJ_HIP(hipSetDevice(c->device));
    c->bgart = 0;
    if (!on) return TJ_OK;
    if ((!c->bg_tmp && dalloc(&c->bg_tmp, c->nvox) != hipSuccess) ||
        (!c->bg_mm && dalloc(&c->bg_mm, 2 * 4096) != hipSuccess) ||
        (!c->bg_hist && dalloc(&c->bg_hist, 256) != hipSuccess)) {
        (void)hipGetLastError();
        return fail(TJ_ENOMEM, "tj_set_bgart: cannot allocate its workspace");
    }
    c->bg_k = k;
    c->bg_dark = dark_bg ? 1 : 0;
    c->bg_median = median ? 1 : 0;
    c->bg_value = bg_value;
    c->bgart = 1;
    // the constructor runs the collapse filter once (:20-27)
    return bgart_collapse(c);
}

// ResolutionEstimation.fsc (:456-557) of two device volumes on `st`: shell
// sums on the device, the four curves of fsc[4][W/2+1] on the host (:545-550).
static int fsc_run(int device, hipStream_t st, const float* even, const float* odd, int W, int H, int D,
                   double* out)
{
    const int R = W / 2 + 1;
    const size_t ws_size = tjk::fsc_workspace(W, H, D);
    void* ws = nullptr;
    double* sums = nullptr;
    if (hipMalloc(&ws, ws_size) != hipSuccess || hipMalloc((void**)&sums, sizeof(double) * 5 * R) != hipSuccess) {
        (void)hipGetLastError();
        if (ws) (void)hipFree(ws);
        return fail(TJ_ENOMEM, "tj_fsc: cannot allocate %zu bytes of workspace", ws_size);
    }
    std::vector<double> h(5 * (size_t)R);
    hipError_t e = tjk::fsc(st, even, odd, W, H, D, ws, ws_size, sums);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), sums, sizeof(double) * 5 * R, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(ws);
    (void)hipFree(sums);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_fsc: %s", hipGetErrorString(e));
    (void)device;
    for (int i = 0; i < R; ++i) {
        const double num = h[i], de = h[R + i], dodd = h[2 * R + i], err = h[3 * R + i], cnt = h[4 * R + i];
        out[i] = (double)i / W;
        out[R + i] = num / std::sqrt(de * dodd);
        out[2 * R + i] = 2 / std::sqrt(cnt);
        out[3 * R + i] = err / cnt;
    }
    return TJ_OK;
}

static int fsc_check_dims(int W, int H, int D)
{
    if (W <= 0 || H <= 0 || D <= 0) return fail(TJ_EINVAL, "tj_fsc: empty volume");
    if (H > 65535 || D > 65535 || sizeof(double) * 5 * (W / 2 + 1) > 64 * 1024)
        return fail(TJ_EINVAL, "tj_fsc: volume %d x %d x %d beyond the supported size", W, H, D);
    return TJ_OK;
}

int tj_fsc(tj_ctx* even, tj_ctx* odd, double* out)
{
    if (!even || !odd || !out) return fail(TJ_EINVAL, "tj_fsc: null argument");
    const TjGeomDev &a = even->g, &b = odd->g;
    if (a.vw != b.vw || a.vh != b.vh || a.vd != b.vd) return fail(TJ_EINVAL, "tj_fsc: volume sizes differ");
    if (even->device != odd->device) return fail(TJ_EINVAL, "tj_fsc: the volumes are on different devices");
    int rc = fsc_check_dims(a.vw, a.vh, a.vd);
    if (rc) return rc;
    TJ_HIP(hipSetDevice(even->device));
    TJ_HIP(hipStreamSynchronize(odd->stream));
    return fsc_run(even->device, even->stream, even->vol, odd->vol, a.vw, a.vh, a.vd, out);
}

int tj_fsc_volumes(int device, const float* even_zyx, const float* odd_zyx, int W, int H, int D, double* out)
{
    if (!even_zyx || !odd_zyx || !out) return fail(TJ_EINVAL, "tj_fsc_volumes: null argument");
    int rc = fsc_check_dims(W, H, D);
    if (rc) return rc;
    TJ_HIP(hipSetDevice(device));
    const size_t n = (size_t)W * H * D;
    float *stage = nullptr, *ev = nullptr, *od = nullptr;
    hipError_t e = hipMalloc((void**)&stage, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&ev, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&od, n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(stage, even_zyx, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = tjk::zyx_to_yzx(nullptr, stage, ev, W, D, H, 0);
    if (e == hipSuccess) e = hipMemcpy(stage, odd_zyx, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = tjk::zyx_to_yzx(nullptr, stage, od, W, D, H, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) rc = fsc_run(device, nullptr, ev, od, W, H, D, out);
    if (stage) (void)hipFree(stage);
    if (ev) (void)hipFree(ev);
    if (od) (void)hipFree(od);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? TJ_ENOMEM : TJ_EDEVICE, "tj_fsc_volumes: %s", hipGetErrorString(e));
    }
    return rc;
}

// ReconstructionThreadManager.getProjection(i) for every image
// (ReconstructionThreadManager.java:257-270): addProjection(i), project(),
// getProjection(0); all n at once into the context's subset buffers.
static int project_all(tj_ctx* c)
{
    int rc = ensure_slots(c, c->nproj);
    if (rc) return rc;
    c->added.resize(c->nproj);
    for (int k = 0; k < c->nproj; ++k) c->added[k] = k;
    if ((rc = upload_added(c))) return rc;
    if ((rc = do_project(c, c->slot_proj, c->added, 0))) return rc;
    c->added.clear();
    c->have_diff = c->have_proj = false;
    return TJ_OK;
}

int tj_ssnr(tj_ctx* sig, tj_ctx* noi, int ring_width, double pixel_size, double* out, int* nfreq, float* vssnr)
{
    if (!sig || !noi || !out || !nfreq) return fail(TJ_EINVAL, "tj_ssnr: null argument");
    const TjGeomDev &a = sig->g, &b = noi->g;
    if (a.pw != b.pw || a.ph != b.ph || sig->nproj != noi->nproj || a.vw != b.vw || a.vh != b.vh || a.vd != b.vd)
        return fail(TJ_EINVAL, "tj_ssnr: the signal and noise sets differ in size");
    if (sig->device != noi->device) return fail(TJ_EINVAL, "tj_ssnr: the sets are on different devices");
    if (sig->nproj <= 0) return fail(TJ_ESTATE, "tj_ssnr: no resident projections");
    if (ring_width < 1) return fail(TJ_EINVAL, "tj_ssnr: ring width %d", ring_width);
    const int W = a.pw, H = a.ph, n = sig->nproj;
    if (sizeof(double) * 5 * W > 64 * 1024) return fail(TJ_EINVAL, "tj_ssnr: detector width %d too large", W);
    TJ_HIP(hipSetDevice(sig->device));
    int rc = project_all(sig);
    if (!rc) rc = project_all(noi);
    if (rc) return rc;
    TJ_HIP(hipStreamSynchronize(noi->stream));
    const size_t ws_size = tjk::ssnr_workspace(W, H);
    void* ws = nullptr;
    double* sums = nullptr;
    float* maps = nullptr;
    if (hipMalloc(&ws, ws_size) != hipSuccess || hipMalloc((void**)&sums, sizeof(double) * 5 * W) != hipSuccess ||
        (vssnr && hipMalloc((void**)&maps, sizeof(float) * sig->npx * n) != hipSuccess)) {
        (void)hipGetLastError();
        if (ws) (void)hipFree(ws);
        if (sums) (void)hipFree(sums);
        return fail(TJ_ENOMEM, "tj_ssnr: cannot allocate %zu bytes of workspace", ws_size);
    }
    std::vector<double> h(5 * (size_t)W);
    hipError_t e = tjk::ssnr(sig->stream, sig->exps, sig->sproj, noi->exps, noi->sproj, W, H, n, ring_width, ws,
                             ws_size, sums, maps);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), sums, sizeof(double) * 5 * W, hipMemcpyDeviceToHost, sig->stream);
    if (e == hipSuccess && maps)
        e = hipMemcpyAsync(vssnr, maps, sizeof(float) * sig->npx * n, hipMemcpyDeviceToHost, sig->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sig->stream);
    (void)hipFree(ws);
    (void)hipFree(sums);
    if (maps) (void)hipFree(maps);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_ssnr: %s", hipGetErrorString(e));
    // SSNR (:212-260): ratios, per-ring means, then rows for the non-negative
    // frequencies; TomoJ divides the dB rows by its counter imgno, which ends
    // at W + 1.
    double *SS = h.data(), *SN = SS + W, *NS = SN + W, *NN = NS + W, *K = NN + W;
    std::vector<double> ssnr1(W), nsnr1(W);
    int imgno = 1;
    for (int i = 0; i < W; ++i) {
        ssnr1[i] = SS[i] / SN[i];
        SS[i] /= K[i];
        SN[i] /= K[i];
        nsnr1[i] = NS[i] / NN[i];
        NS[i] /= K[i];
        NN[i] /= K[i];
        imgno++;
    }
    const int R = W / 2 + 1;
    int imax = 0;
    for (int i = 0; i < W; ++i) {
        const double w = (double)((i <= (W >> 1)) ? i : i - W) / W;  // convertToFrequency
        if (w < 0) {
            imax = i;
            break;
        }
        const double ssnr = ssnr1[i] / nsnr1[i];
        out[0 * R + i] = i;
        out[1 * R + i] = w / pixel_size;
        out[2 * R + i] = ssnr > 1 ? 10 * std::log10(ssnr - 1) : -1000;
        out[3 * R + i] = ssnr1[i];
        out[4 * R + i] = 10 * std::log10(SS[i] / imgno);
        out[5 * R + i] = 10 * std::log10(SN[i] / imgno);
        out[6 * R + i] = nsnr1[i];
        out[7 * R + i] = 10 * std::log10(NS[i] / imgno);
        out[8 * R + i] = 10 * std::log10(NN[i] / imgno);
        imax++;
    }
    *nfreq = imax;
    return TJ_OK;
}

int tj_available_indexes(int nproj, int projection_type, int* out)
{
    if (nproj < 0 || (nproj > 0 && !out)) return fail(TJ_EINVAL, "tj_available_indexes: bad arguments");
    // ReconstructionParameters.getAvailableIndexes (:212-256): EVEN 11 / ODD 12
    // keep that parity, at most nproj/2 images; any other type keeps all.
    std::vector<char> take(nproj);
    int cnt = 0;
    for (int i = 0; i < nproj; ++i) {
        if (projection_type == 11 || projection_type == 12) {
            const bool par = (projection_type == 11) ? (i % 2 == 0) : (i % 2 == 1);
            take[i] = par;
            if (par) ++cnt;
            if (cnt > nproj / 2) {
                take[i] = 0;
                --cnt;
            }
        } else {
            take[i] = 1;
            ++cnt;
        }
    }
    int k = 0;
    for (int i = 0; i < nproj; ++i)
        if (take[i]) out[k++] = i;
    return cnt;
}

int tj_set_available_indexes(tj_ctx* c, const int* indexes, int n)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_available_indexes: null context");
    if (n < 0 || (n > 0 && !indexes)) return fail(TJ_EINVAL, "tj_set_available_indexes: bad list");
    c->avail.assign(indexes, indexes + n);
    return TJ_OK;
}

int tj_set_bayesian(tj_ctx* c, int on, double alpha, double lambda)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_bayesian: null context");
    if (on && (c->tvm_iter > 0 || c->bgart)) return fail(TJ_ESTATE, "tj_set_bayesian: another regulariser is on");
    TJ_HIP(hipSetDevice(c->device));
    if (on) {
        int rc = ensure_dist(c);
        if (rc) return rc;
    }
    c->bayes = on ? 1 : 0;
    c->bayes_alpha = alpha;
    c->bayes_lambda = lambda;
    return TJ_OK;
}

int tj_set_tvm(tj_ctx* c, int iterations, double theta, double g, double lambda)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_tvm: null context");
    if (iterations < 0) return fail(TJ_EINVAL, "tj_set_tvm: negative iteration count %d", iterations);
    if (iterations > 0 && (c->bayes || c->bgart)) return fail(TJ_ESTATE, "tj_set_tvm: another regulariser is on");
    TJ_HIP(hipSetDevice(c->device));
    if (iterations > 0) {
        int rc = ensure_tv(c);
        if (rc) return rc;
    }
    c->tvm_iter = iterations;
    c->tvm_theta = (float)theta;  // setTvmTheta / setTvmG cast to float (:64-74)
    c->tvm_g = (float)g;
    c->tvm_dt = (float)lambda;    // dt = (float) regularizationLambda (:152)
    return TJ_OK;
}

int tj_tv_denoise(tj_ctx* c, float theta, float g, float dt, int iterations)
{
    if (!c) return fail(TJ_EINVAL, "tj_tv_denoise: null context");
    if (iterations < 0) return fail(TJ_EINVAL, "tj_tv_denoise: negative iteration count %d", iterations);
    TJ_HIP(hipSetDevice(c->device));
    return tv_run(c, theta, g, dt, iterations);
}

int tj_set_fista(tj_ctx* c, int on)
{
    if (!c) return fail(TJ_EINVAL, "tj_set_fista: null context");
    TJ_HIP(hipSetDevice(c->device));
    if (on && !c->prev) {
        if (dalloc(&c->prev, c->nvox) != hipSuccess) {
            (void)hipGetLastError();
            return fail(TJ_ENOMEM, "tj_set_fista: cannot allocate the previous-iteration volume");
        }
    }
    c->fista = on ? 1 : 0;
    c->fista_t = 0.0;  // FistaProjector3D.java:8
    return TJ_OK;
}

int tj_start_iteration(tj_ctx* c)
{
    if (!c) return fail(TJ_EINVAL, "null context");
    TJ_HIP(hipSetDevice(c->device));
    int rc = start_hooks(c);
    if (rc) return rc;
    TJ_HIP(hipStreamSynchronize(c->stream));
    return TJ_OK;
}

int tj_end_iteration(tj_ctx* c)
{
    if (!c) return fail(TJ_EINVAL, "null context");
    TJ_HIP(hipSetDevice(c->device));
    return end_hooks(c);
}

double tj_get_completion(tj_ctx* c) { return c ? c->completion.load() : 0.0; }

int tj_interrupt(tj_ctx* c)
{
    if (!c) return fail(TJ_EINVAL, "tj_interrupt: null context");
    c->interrupted.store(1);
    return TJ_OK;
}

int tj_enable_timing(tj_ctx* c, int on)
{
    if (!c) return fail(TJ_EINVAL, "tj_enable_timing: null context");
    c->timing_on = on ? 1 : 0;
    return TJ_OK;
}

int tj_get_timing(tj_ctx* c, tj_timing* t)
{
    if (!c || !t) return fail(TJ_EINVAL, "tj_get_timing: null argument");
    *t = c->last_timing;
    return TJ_OK;
}

// TomoReconstruction2.OSSART + osartIteration (:548-733) over the resident
// projections; all device work is queued on the context stream and the host
// synchronises once per iteration to read the errors back.
int tj_run_ossart(tj_ctx* c, int iterations, double lambda, int update_nb, int first_index, int start_y, int end_y,
                  double* err)
{
    if (!c) return fail(TJ_EINVAL, "tj_run_ossart: null context");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    const int n = c->nproj;
    if (n <= 0) return fail(TJ_ESTATE, "tj_run_ossart: no resident projections");
    if (iterations < 0 || update_nb < 0) return fail(TJ_EINVAL, "tj_run_ossart: negative count");
    if ((rc = check_available(c, "tj_run_ossart"))) return rc;
    const int nav = num_available(c);
    if (first_index < 0 || (nav > 0 && first_index >= nav))
        return fail(TJ_EINVAL, "tj_run_ossart: first index %d of %d", first_index, nav);
    TJ_HIP(hipSetDevice(c->device));
    const int update = (update_nb == 0) ? n : update_nb;
    if ((rc = ensure_slots(c, std::min(update, n)))) return rc;
    if ((rc = ensure_iota(c, n))) return rc;
    c->interrupted.store(0);
    c->added.clear();
    c->have_diff = c->have_proj = false;

    std::vector<int> order;
    order_shuffle(c, first_index, order);
    if (nav > 0) TJ_HIP(hipMemcpy(c->order, order.data(), sizeof(int) * nav, hipMemcpyHostToDevice));
    const float factorimg = (float)(lambda / update);
    const double completion_inc = (end_y - start_y) / (double)c->g.vh;

    tj_timing tm{};
    size_t nev = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (c->timing_on) {
        t0 = ev(c, nev++);
        t1 = ev(c, nev++);
        TJ_HIP(hipMemsetAsync(c->bpslow + 1, 0, sizeof(int), c->stream));
        TJ_HIP(hipEventRecord(t0, c->stream));
    }
    std::vector<std::pair<hipEvent_t, hipEvent_t>> fpev, resev, prepev, bpev;
    for (int it = 0; it < iterations; ++it) {
        if (c->interrupted.load()) return fail(TJ_ECANCELLED, "tj_run_ossart: interrupted");
        // projector.startOfIteration() / endOfIteration() around each
        // osartIteration (TomoReconstruction2.java:593-595)
        if ((rc = start_hooks(c))) return rc;
        // With an available list the loop walks it, but update (0 -> n), the
        // flush test t == n-1 and the error normalisation keep the full count
        // n (:639-644, :662, :726): a last partial subset is never projected.
        int diffIndex = 0, done = 0;
        for (int t = 0; t < nav; ++t) {
            ++diffIndex;
            if (diffIndex == update || t == n - 1) {
                const int off = t - diffIndex + 1;
                const int* slots = c->order + off;
                std::vector<int> hs(order.begin() + off, order.begin() + off + diffIndex);
                hipEvent_t a = nullptr, b = nullptr, r = nullptr, q = nullptr, d = nullptr;
                if (c->timing_on) {
                    a = ev(c, nev++);
                    b = ev(c, nev++);
                    r = ev(c, nev++);
                    q = ev(c, nev++);
                    d = ev(c, nev++);
                    TJ_HIP(hipEventRecord(a, c->stream));
                }
                const int start = fp_start_row(c, start_y);
                const int fk = pick_fp_kernel(c, hs, start);
                const int bk = pick_bp_kernel(c, hs);
                bool slow_cleared = false;  // the fold zeroed the backprojection's row flag
                if (fk == 3 && !c->geo.long_object_comp) {
                    // projectAndDifference in one pass: the LDS forward kernel
                    // writes the differences and the row errors itself
                    ++tm.fused_subsets;
                    TJ_HIP(tjk::fp_residual(c->stream, c->g, c->vol, c->pp, slots, diffIndex, start, c->norm_tab,
                                            c->tune.fp_rays, c->tune.fp_zchunk, c->exps, (double)factorimg,
                                            c->sdiff, c->row_err, c->rbad, c->rowflag, c->err + off, b,
                                            c->tune.fp_groups, c->bpslow, c->tune.fp_prefetch == 0));
                    slow_cleared = true;
                    if (c->timing_on) TJ_HIP(hipEventRecord(r, c->stream));
                } else {
                    TJ_HIP(tjk::fp(c->stream, c->g, c->vol, c->pp, slots, diffIndex, start, fk, c->pflag,
                                   c->norm_tab, c->sproj, c->snorm, c->tune.fp_rows, c->tune.fp_rays,
                                   c->tune.fp_zchunk, c->tune.fp_groups, c->tune.fp_prefetch == 0));
                    if (c->geo.long_object_comp)
                        TJ_HIP(tjk::norm_max_fill(c->stream, (int)c->npx, diffIndex, c->snorm));
                    if (c->timing_on) TJ_HIP(hipEventRecord(b, c->stream));
                    if ((rc = do_residual(c, slots, diffIndex, (double)factorimg, c->err + off))) return rc;
                    if (c->timing_on) TJ_HIP(hipEventRecord(r, c->stream));
                }
                if (bk == 3)
                    if ((rc = ensure_bpairs(c, diffIndex))) return rc;
                TJ_HIP(tjk::bp(c->stream, c->g, bp_target(c), c->pp, slots, diffIndex, c->sdiff, c->iota,
                               c->rowflag, c->bprows, start_y, end_y, bk, c->tune.bp_rows, c->tune.bp_zchunk,
                               c->tune.bp_group, bp_finite(c), c->bpairs, q, c->bpslow, c->tune.bp_fast,
                               slow_cleared ? 1 : 0));
                if ((rc = do_error_backproject(c, slots, hs, start_y, end_y))) return rc;
                if (c->timing_on) {
                    TJ_HIP(hipEventRecord(d, c->stream));
                    fpev.emplace_back(a, b);
                    resev.emplace_back(b, r);
                    prepev.emplace_back(r, q);
                    bpev.emplace_back(q, d);
                }
                done = t + 1;
                diffIndex = 0;
                if (c->interrupted.load()) {
                    (void)hipStreamSynchronize(c->stream);
                    return fail(TJ_ECANCELLED, "tj_run_ossart: interrupted");
                }
            }
        }
        if (done > 0)
            TJ_HIP(hipMemcpyAsync(c->h_err, c->err, sizeof(double) * done, hipMemcpyDeviceToHost, c->stream));
        TJ_HIP(hipStreamSynchronize(c->stream));
        // error statistics, in the order osartIteration accumulates them (:692-701, :726-730)
        double som = 0, errmin = 3.4028234663852886e38, errmax = 0;
        int imin = 0, imax = 0;
        for (int t = 0; t < done; ++t) {
            const double e = c->h_err[t];
            som += e;
            if (e > errmax) {
                errmax = e;
                imax = order[t];
            } else if (e < errmin) {
                errmin = e;
                imin = order[t];
            }
        }
        if (err) {
            const double wh = (double)c->g.vw * c->g.vh;
            double* e5 = err + 5 * it;
            e5[0] = std::sqrt(som / (double)((long)n * c->g.vw * c->g.vh));
            e5[1] = std::sqrt(errmin / wh);
            e5[2] = imin;
            e5[3] = std::sqrt(errmax / wh);
            e5[4] = imax;
        }
        if ((rc = end_hooks(c))) return rc;
        c->completion.store(c->completion.load() + completion_inc);
    }
    if (c->timing_on) {
        TJ_HIP(hipEventRecord(t1, c->stream));
        TJ_HIP(hipEventSynchronize(t1));
        float ms = 0;
        TJ_HIP(hipEventElapsedTime(&ms, t0, t1));
        tm.total_ms = ms;
        for (auto& p : fpev) {
            TJ_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            tm.fp_ms += ms;
            tm.fp_launches++;
        }
        for (auto& p : resev) {
            TJ_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            tm.res_ms += ms;
            tm.res_launches++;
        }
        for (auto& p : prepev) {
            TJ_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            tm.bpprep_ms += ms;
            tm.bpprep_launches++;
        }
        for (auto& p : bpev) {
            TJ_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            tm.bp_ms += ms;
            tm.bp_launches++;
        }
        tm.other_ms = tm.total_ms - tm.fp_ms - tm.res_ms - tm.bpprep_ms - tm.bp_ms;
        int nfast = 0;
        TJ_HIP(hipMemcpy(&nfast, c->bpslow + 1, sizeof(int), hipMemcpyDeviceToHost));
        tm.fast_bp_launches = nfast;
    }
    c->last_timing = tm;
    c->completion.store(0.0);
    return TJ_OK;
}

// TomoReconstruction2.WBP (:499-538) with update = 0: every resident projection
// is weighted (when a filter is set) and backprojected in index order.  The
// per-voxel accumulation order, with its scrub after each image, is the same as
// backprojecting one image per call.
int tj_run_wbp(tj_ctx* c, int start_y, int end_y)
{
    if (!c) return fail(TJ_EINVAL, "tj_run_wbp: null context");
    int rc = check_range(c, start_y, end_y);
    if (rc) return rc;
    const int n = c->nproj;
    if (n <= 0) return fail(TJ_ESTATE, "tj_run_wbp: no resident projections");
    // the Bayesian projector's backproject adds nothing without differences (:90)
    if (c->bayes) return TJ_OK;
    TJ_HIP(hipSetDevice(c->device));
    if ((rc = check_available(c, "tj_run_wbp"))) return rc;
    if ((rc = ensure_slots(c, n))) return rc;
    if ((rc = ensure_iota(c, n))) return rc;
    c->added.clear();
    c->have_diff = c->have_proj = false;
    // TomoReconstruction2.WBP walks getAvailableIndexes (:500)
    std::vector<int> idx(num_available(c));
    for (int k = 0; k < (int)idx.size(); ++k) idx[k] = c->avail.empty() ? k : c->avail[k];
    if (idx.empty()) return TJ_OK;
    hipEvent_t t0 = nullptr, t1 = nullptr, tb = nullptr;
    if (c->timing_on) {
        t0 = ev(c, 0);
        tb = ev(c, 1);
        t1 = ev(c, 2);
        TJ_HIP(hipEventRecord(t0, c->stream));
    }
    const int* slots = c->iota;
    if (!c->avail.empty()) {
        TJ_HIP(hipMemcpy(c->slot_proj, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
        c->uploaded.clear();
        slots = c->slot_proj;
    }
    const float* images;
    const int* index;
    if ((rc = prepare_originals(c, slots, idx, &images, &index))) return rc;
    if (c->timing_on) TJ_HIP(hipEventRecord(tb, c->stream));
    if ((rc = do_backproject(c, slots, idx, images, index, start_y, end_y))) return rc;
    if (c->timing_on) {
        TJ_HIP(hipEventRecord(t1, c->stream));
        TJ_HIP(hipEventSynchronize(t1));
        float a = 0, b = 0;
        TJ_HIP(hipEventElapsedTime(&a, t0, t1));
        TJ_HIP(hipEventElapsedTime(&b, tb, t1));
        tj_timing tm{};
        tm.total_ms = a;
        tm.bp_ms = b;
        tm.bp_launches = 1;
        tm.other_ms = a - b;
        c->last_timing = tm;
    }
    TJ_HIP(hipStreamSynchronize(c->stream));
    return TJ_OK;
}

// ---------------------------------------------------------------------------
// synthetic inputs
// ---------------------------------------------------------------------------
static inline uint64_t splitmix64_next(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t v) { return (double)(v >> 11) * 0x1.0p-53; }

void tj_phantom_ellipsoids(uint64_t seed, int k, int W, int H, int D, double* out)
{
    uint64_t s = seed;
    double L = std::min(std::min(W, H), D);
    for (int e = 0; e < k; ++e) {
        double* o = out + 16 * e;
        o[0] = 0.1 + 0.9 * u01(splitmix64_next(&s));
        double px, py, pz;
        do {
            px = 2 * u01(splitmix64_next(&s)) - 1;
            py = 2 * u01(splitmix64_next(&s)) - 1;
            pz = 2 * u01(splitmix64_next(&s)) - 1;
        } while (px * px + py * py + pz * pz > 1.0);
        o[1] = px * 0.4 * W;
        o[2] = py * 0.4 * H;
        o[3] = pz * 0.4 * D;
        for (int a = 0; a < 3; ++a) o[4 + a] = (0.03 + 0.22 * u01(splitmix64_next(&s))) * L;
        const double u1 = u01(splitmix64_next(&s)), u2 = u01(splitmix64_next(&s)), u3 = u01(splitmix64_next(&s));
        const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
        const double qw = a * std::sin(2 * M_PI * u2), qx = a * std::cos(2 * M_PI * u2);
        const double qy = b * std::sin(2 * M_PI * u3), qz = b * std::cos(2 * M_PI * u3);
        double* R = o + 7;
        R[0] = 1 - 2 * (qy * qy + qz * qz);
        R[1] = 2 * (qx * qy - qz * qw);
        R[2] = 2 * (qx * qz + qy * qw);
        R[3] = 2 * (qx * qy + qz * qw);
        R[4] = 1 - 2 * (qx * qx + qz * qz);
        R[5] = 2 * (qy * qz - qx * qw);
        R[6] = 2 * (qx * qz - qy * qw);
        R[7] = 2 * (qy * qz + qx * qw);
        R[8] = 1 - 2 * (qx * qx + qy * qy);
    }
}

int tj_synth_projections(tj_ctx* c, const double* ell, int k, const double* euler, int n, double noise_rel,
                         uint64_t noise_seed)
{
    if (!c || !ell || !euler || n <= 0 || k < 0) return fail(TJ_EINVAL, "tj_synth_projections: bad argument");
    TJ_HIP(hipSetDevice(c->device));
    int rc = alloc_projections(c, n, euler);
    if (rc) return rc;
    double *dell = nullptr, *clean = nullptr, *part = nullptr;
    const int npart = 256;
    const size_t total = c->npx * n;
    if (dalloc(&dell, (size_t)16 * std::max(k, 1)) || dalloc(&clean, total) || dalloc(&part, npart)) {
        (void)hipGetLastError();
        dfree(dell);
        dfree(clean);
        dfree(part);
        return fail(TJ_ENOMEM, "tj_synth_projections: cannot allocate %zu doubles", total);
    }
    hipError_t e = hipMemcpy(dell, ell, sizeof(double) * 16 * k, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = tjk::phantom_project(c->stream, dell, k, c->pp, n, c->g.pw, c->g.ph, c->g.pc0, c->g.pc1, clean);
    if (e == hipSuccess) e = tjk::add_noise(c->stream, clean, total, noise_seed, noise_rel, part, npart, c->exps);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(dell);
    dfree(clean);
    dfree(part);
    if (e != hipSuccess) return fail(TJ_EDEVICE, "tj_synth_projections: %s", hipGetErrorString(e));
    return TJ_OK;
}

}  // extern "C"
