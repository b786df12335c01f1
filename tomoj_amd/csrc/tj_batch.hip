// tj_batch.hip -- the job engine: many independent SIRT reconstructions of
// tilt series around a vertical axis, resident on one GPU and advanced by
// persistent kernels that walk one table of tiles over all the jobs of a run.
//
// The arithmetic is TomoJ's algorithm (VoxelProjector3D.projectJosephPartial,
// difference and backProjectPartial; TomoReconstruction2.osartIteration with
// update = N), evaluated with closed-form positions instead of Java's
// incremental double sums:
//   forward  x(i, z) = fma(z, tan, base_i)   base_i = i/cos + b   (x of slice z)
//   back     xx(i, k) = fma(k, e02, xe_i)    xe_i = i*e00 + b2
// in the storage's arithmetic type A (float for FP32/BF16/FP16 storage, double
// for FP64 storage).  Sampling rules are TomoJ's: getPixel's nearest value on
// the last column, the last row and the last slice (TomoReconstruction2.java:154-190), the
// entry/exit norm of the ray (VoxelProjector3D.java:756-805), ImageJ's bilinear
// value with the last column clamped in the backprojection, positivity after
// each image.  Deviations, all of which vanish for finite data: the NaN/Inf
// scrub of the backprojection runs once per voxel after all images instead of
// after each image, and the ray norm is count*normIncrement rounded once.
//
// Layouts per job: volume T[y][z][x], measured projections T[n][ph][pw],
// differences A[n][ph][pw], error partials double[n][ph][column tiles].
#include "tj_device.h"
#include "../../include/tomoj_hip.h"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <string>
#include <vector>

namespace tjk {
hipError_t phantom_project(hipStream_t, const double*, int, const TjProj*, int, int, int, double, double, double*);
hipError_t add_noise(hipStream_t, const double*, size_t, uint64_t, double, double*, int, float*);
}  // namespace tjk

namespace tjb {

constexpr int NT = 256;  // threads per block: one column (forward) or voxel column (back) each
constexpr int PRM = 10;  // doubles of per-projection constants
constexpr int SU = 4;    // columns per lane per staging batch of the forward projection

struct Bf16 {
    uint16_t u;
};

// ---------------------------------------------------------------------------
// storage conversions (round to nearest even on store)
// ---------------------------------------------------------------------------
// Every buffer of a job lives in device global memory: accesses go through
// global (address space 1) pointers so the compiler emits global rather than
// flat instructions (flat loads also count against the LDS wait counter).
#define TJ_GLOBAL __attribute__((address_space(1)))
#define TJ_LDS __attribute__((address_space(3)))
// LDS addresses are 32-bit byte offsets computed in registers and turned into
// address-space-3 pointers where they are read; the host pass, which sees
// 64-bit pointers, would warn about each of those casts.
#pragma clang diagnostic ignored "-Wint-to-pointer-cast"
template <typename T>
__device__ __forceinline__ const TJ_GLOBAL T* gp(const T* p) { return (const TJ_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ TJ_GLOBAL T* gp(T* p) { return (TJ_GLOBAL T*)p; }

template <typename A>
__device__ __forceinline__ A ld(const float* p, size_t i) { return (A)gp(p)[i]; }
template <typename A>
__device__ __forceinline__ A ld(const double* p, size_t i) { return (A)gp(p)[i]; }
template <typename A>
__device__ __forceinline__ A ld(const _Float16* p, size_t i) { return (A)(float)gp(p)[i]; }
template <typename A>
__device__ __forceinline__ A ld(const Bf16* p, size_t i) { return (A)__uint_as_float((uint32_t)gp(p)[i].u << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)(((b >> 16) & 0x8000u) | 0x7fc0u);  // NaN
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ void st(float* p, size_t i, float v) { gp(p)[i] = v; }
__device__ __forceinline__ void st(double* p, size_t i, double v) { gp(p)[i] = v; }
__device__ __forceinline__ void st(_Float16* p, size_t i, float v) { gp(p)[i] = (_Float16)v; }
__device__ __forceinline__ void st(Bf16* p, size_t i, float v) { gp(p)[i].u = f32_to_bf16(v); }

template <typename A> struct V4;
template <> struct V4<float> { using t = float4; };
template <> struct V4<double> { using t = double4; };
template <typename A> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Device descriptor of one job (host fills it; see tj_batch_add_job).
struct TjJob {
    void* vol;
    const void* exps;
    void* diff;
    double* err_part;
    const double* prm;   // [n][PRM]: a, b, tan, normInc, e00, e02, b2, 1/tan, 1/e02
    float* master;       // FP32 master volume [y][z][x] when the stored volume is 16-bit (or null)
    uint32_t* pairs;     // 16-bit storage: differences as FP16 pairs (d[x], d[x+1] - d[x]), [n][ph][pw]
    const int* order;    // [n] order in which the backprojection adds the images
    double* err_hist;    // [hist_cap] sum of diff^2 per iteration
    int* iter;           // iterations done (device counter)
    int W, H, D, pw, ph, n;
    int dy;              // forward: volume row of detector row j is j + dy
    int dyb;             // back: image row of volume row j is j + dyb
    int nct;             // forward column tiles
    int hist_cap;
    double factor;       // (double)(float)(lambda / n)
};

// ---------------------------------------------------------------------------
// Forward projection fused with the residual.  Tile (job, first projection,
// block of 4 detector rows, 256-column tile): the 4 volume rows those detector
// rows see are staged in LDS one slab of zc slices at a time, one A4 cell per
// (z, x) plus a replicated last column and a zero cell; each thread follows
// the rays of its column for G projections.
// ---------------------------------------------------------------------------
// Integers z in [0, n) where the line x(z) = fma(z, t, b), evaluated in double,
// satisfies lo <= x(z) < hi: returned as [zlo, zhi] (empty when zhi < zlo).
// x is monotone in z, so the set is an interval; its ends are estimated from
// the crossings and then settled by evaluating the predicate itself, so every
// arithmetic type decides inside/outside exactly as the FP64 evaluation does.
__device__ inline void line_interval(double b, double t, double inv_t, double lo, double hi, int n, int& zlo,
                                     int& zhi)
{
    auto inside = [&](int z) {
        const double x = fma((double)z, t, b);
        return x >= lo && x < hi;
    };
    int a, c;
    if (t == 0.0) {
        a = 0;
        c = inside(0) ? n - 1 : -1;
    } else {
        const double za = (lo - b) * inv_t, zb = (hi - b) * inv_t;
        const double zmin = fmin(za, zb), zmax = fmax(za, zb);
        a = (int)fmin(fmax(ceil(zmin), 0.0), (double)(n - 1));
        c = (int)fmin(fmax(floor(zmax), 0.0), (double)(n - 1));
        for (int s = 0; s < 3 && a > 0 && inside(a - 1); ++s) --a;
        for (int s = 0; s < 3 && a <= c && !inside(a); ++s) ++a;
        for (int s = 0; s < 3 && c < n - 1 && inside(c + 1); ++s) ++c;
        for (int s = 0; s < 3 && c >= a && !inside(c); ++s) --c;
    }
    zlo = a;
    zhi = c;
}

// Fraction of a non-negative position xs with integer part xb.
__device__ __forceinline__ float frac_of(float xs, int) { return __builtin_amdgcn_fractf(xs); }
__device__ __forceinline__ double frac_of(double xs, int xb) { return xs - (double)xb; }

// Adds the samples of one slice for the G rays of the thread: q0/q1 are the
// cells at and right of each ray's column, xf its x fraction.
template <typename A, int G, bool NEARY>
__device__ __forceinline__ void fp_accumulate(const typename V4<A>::t (&q0)[G], const typename V4<A>::t (&q1)[G],
                                              const A (&xf)[G], const A (&ym)[4], A (&rs)[G][4])
{
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (NEARY) {
            rs[k][0] += fma(xf[k] * ym[0], q1[k].x - q0[k].x, q0[k].x);
            rs[k][1] += fma(xf[k] * ym[1], q1[k].y - q0[k].y, q0[k].y);
            rs[k][2] += fma(xf[k] * ym[2], q1[k].z - q0[k].z, q0[k].z);
            rs[k][3] += fma(xf[k] * ym[3], q1[k].w - q0[k].w, q0[k].w);
        } else {
            rs[k][0] += fma(xf[k], q1[k].x - q0[k].x, q0[k].x);
            rs[k][1] += fma(xf[k], q1[k].y - q0[k].y, q0[k].y);
            rs[k][2] += fma(xf[k], q1[k].z - q0[k].z, q0[k].z);
            rs[k][3] += fma(xf[k], q1[k].w - q0[k].w, q0[k].w);
        }
    }
}

// Slice z of the G rays with FP64 positions (nearest-value samples: the x
// column of a discontinuous sample must be the one TomoJ's FP64 position
// gives); zl = 0 on the last slice.
template <typename A, int G, bool NEARY>
__device__ __forceinline__ void fp_slice_f64(const typename V4<A>::t* cell, int z, int zo, int zero, A zl,
                                             double wd, const A (&ym)[4], const double (&bd)[G],
                                             const double (&td)[G], const int (&zlo)[G], const int (&zhi)[G],
                                             A (&rs)[G][4])
{
    using C4 = typename V4<A>::t;
    int a[G];
    A xf[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const double xv = fma((double)z, td[k], bd[k]);
        const bool in = (z >= zlo[k]) && (z <= zhi[k]);
        const double xs = fmin(fmax(xv, 0.0), wd);
        const int xb = (int)xs;
        xf[k] = (A)(xs - (double)xb) * zl;
        a[k] = in ? zo + xb : zero;
    }
    C4 q0[G], q1[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        q0[k] = cell[a[k]];
        q1[k] = cell[a[k] + 1];
    }
    fp_accumulate<A, G, NEARY>(q0, q1, xf, ym, rs);
}

// x clamped to [0, hi] (hi >= 0, x finite): one v_med3 in FP32.
template <typename A>
__device__ __forceinline__ A clamp0(A x, A hi)
{
    return fmin(fmax(x, (A)0), hi);
}
template <>
__device__ __forceinline__ float clamp0<float>(float x, float hi)
{
    return __builtin_amdgcn_fmed3f(x, 0.0f, hi);
}

// The A4 cell at LDS byte address a.
template <typename A>
__device__ __forceinline__ typename V4<A>::t lds_cell(uint32_t a)
{
    typedef A v4 __attribute__((ext_vector_type(4)));
    const v4 t = *(const TJ_LDS v4*)a;
    typename V4<A>::t c;
    c.x = t.x;
    c.y = t.y;
    c.z = t.z;
    c.w = t.w;
    return c;
}

// The FP64 ray constants (x = z * td + bd) of the thread's G rays, from the
// job's parameter table (column ic, projections p0 .. p0 + G - 1).
template <int G>
__device__ __forceinline__ void ray_f64(const double* prm, int p0, int n, double ic, double (&bd)[G], double (&td)[G])
{
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const double* q = prm + PRM * min(p0 + k, n - 1);
        bd[k] = ic * q[0] + q[1];
        td[k] = q[2];
    }
}

// One LDS slab of the forward projection: zn slices for the G rays of the
// thread.  Interpolated samples take their position in the arithmetic type A
// (the bilinear value is continuous in x, so rounding moves it by a rounding
// error).  getPixel's nearest-value samples -- the last slice (zlast) and, with
// NEARY, the last volume row (ym = 0) -- are discontinuous in x, so there the
// column comes from the FP64 position, as TomoJ computes it: f64(neary, z, zo,
// zl) projects such a slice (fp_slice_f64 with the rays' FP64 constants).  The
// main loop has no branch: addresses, then LDS reads, then arithmetic, for all
// G rays at once, or with BYRAY ray by ray (fewer registers live: 2 cells of
// one ray instead of 2 G).
template <typename A, int G, bool NEARY, typename F64, bool BYRAY = false>
__device__ __forceinline__ void fp_slab(const typename V4<A>::t* cell, int zn, int zlast, int z0, int W1, int zero,
                                        A wm1, const A (&ym)[4], const A (&base)[G], const A (&tn)[G],
                                        const int (&zlo)[G], const int (&zhi)[G], A (&rs)[G][4], F64 f64)
{
    using C4 = typename V4<A>::t;
    const int zend = min(zlast, zn);
    if (NEARY) {
        for (int zz = 0; zz < zend; ++zz) f64(std::true_type{}, z0 + zz, zz * W1, (A)1);
    } else {
        for (int zz = 0; zz < zend; ++zz) {
            const int z = z0 + zz;
            const int zo = zz * W1;
            const A zf = (A)z;
            // LDS byte address of the slice's first cell, kept in a scalar
            // register, so a cell address is one shift-add of its index; the
            // zero cell counted from the slice
            uint32_t sz = (uint32_t)(size_t)(const TJ_LDS C4*)(cell + zo);
            asm("" : "+s"(sz));
            const int zrel = zero - zo;
            uint32_t a[G];
            A xf[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                // inside the volume (decided in FP64 by zlo/zhi) xv lies in
                // [-ACCURACY, W) up to rounding: x < 0 samples column 0 with
                // fraction 0 (getPixel), x >= W-1 the last column with fraction
                // 0 (its replicated neighbour makes the fraction irrelevant).
                // Outside, the lane reads the zero cell.
                const A xs = clamp0(fma(zf, tn[k], base[k]), wm1);
                const int xb = (int)xs;
                xf[k] = frac_of(xs, xb);
                const int c = ((z >= zlo[k]) && (z <= zhi[k])) ? xb : zrel;
                a[k] = sz + (uint32_t)sizeof(C4) * (uint32_t)c;
            }
            if constexpr (BYRAY) {
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const C4 q0 = lds_cell<A>(a[k]), q1 = lds_cell<A>(a[k] + (uint32_t)sizeof(C4));
                    rs[k][0] += fma(xf[k], q1.x - q0.x, q0.x);
                    rs[k][1] += fma(xf[k], q1.y - q0.y, q0.y);
                    rs[k][2] += fma(xf[k], q1.z - q0.z, q0.z);
                    rs[k][3] += fma(xf[k], q1.w - q0.w, q0.w);
                }
            } else {
                C4 q0[G], q1[G];
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    q0[k] = lds_cell<A>(a[k]);
                    q1[k] = lds_cell<A>(a[k] + (uint32_t)sizeof(C4));
                }
                fp_accumulate<A, G, false>(q0, q1, xf, ym, rs);
            }
        }
    }
    if (zlast < zn) f64(std::integral_constant<bool, NEARY>{}, z0 + zlast, zlast * W1, (A)0);
}

// One element of a volume row read through a buffer resource (element index
// within the row's plane; a row outside the volume has an empty range and
// reads 0), widened to the arithmetic type.
template <typename A>
__device__ __forceinline__ A bld(const float*, __amdgpu_buffer_rsrc_t rs, uint32_t e)
{
    return (A)__uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, e * 4u, 0, 0));
}
template <typename A>
__device__ __forceinline__ A bld(const double*, __amdgpu_buffer_rsrc_t rs, uint32_t e)
{
    return (A)__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, e * 8u, 0, 0));
}
template <typename A>
__device__ __forceinline__ A bld(const _Float16*, __amdgpu_buffer_rsrc_t rs, uint32_t e)
{
    return (A)(float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(rs, e * 2u, 0, 0));
}
template <typename A>
__device__ __forceinline__ A bld(const Bf16*, __amdgpu_buffer_rsrc_t rs, uint32_t e)
{
    return (A)__uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rs, e * 2u, 0, 0) << 16);
}

// At most 128 VGPRs (4 waves per SIMD): the kernel is latency bound and a CU
// holding four of its blocks hides one block's slab staging under the others'
// projection.
template <typename T, typename TE, typename A, int G>
__global__ __launch_bounds__(NT, 4) void k_job_fp(const TjJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                                  int ntiles, int lds_cells)
{
    using C4 = typename V4<A>::t;
    extern __shared__ __align__(16) unsigned char smem[];
    C4* cell = reinterpret_cast<C4*>(smem);
    __shared__ double red[NT / 64][G * 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        if (tl.x < 0) continue;  // padding entry (uniform in the block)
        const TjJob& J = jobs[tl.x];
        const T* vol = (const T*)J.vol;
        const int W = J.W, D = J.D, pw = J.pw, ph = J.ph, n = J.n;
        const int W1 = W + 1;
        const int zc = min(D, (lds_cells - 2) / W1);
        const int zero = zc * W1;
        const int p0 = tl.y, j0 = tl.z * 4, ct = tl.w;
        const int i = ct * NT + tid;
        const bool col = i < pw;
        const double ic = col ? i : pw - 1;
        const size_t plane = (size_t)D * W;
        bool rin[4];
        __amdgpu_buffer_rsrc_t rsc[4];  // the 4 volume rows; empty outside the volume
        A ym[4];  // 0 on the last volume row: getPixel takes the nearest value there
        bool neary = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int y = j0 + r + J.dy;
            rin[r] = (j0 + r < ph) && y >= 0 && y < J.H;
            rsc[r] = __builtin_amdgcn_make_buffer_rsrc((void*)(vol + (size_t)(rin[r] ? y : 0) * plane), 0,
                                                       rin[r] ? (int)(plane * sizeof(T)) : 0, 0x00020000);
            ym[r] = (y == J.H - 1) ? (A)0 : (A)1;
            neary = neary || (y == J.H - 1);
        }
        A base[G], tn[G], rs[G][4];
        double bd[G], td[G];
        int zlo[G], zhi[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const double* q = J.prm + PRM * min(p0 + k, n - 1);
            bd[k] = ic * q[0] + q[1];
            td[k] = q[2];
            base[k] = (A)bd[k];
            tn[k] = (A)td[k];
            // inside test of VoxelProjector3D.java:761 (x + ACCURACY >= 0, x < W)
            line_interval(bd[k], td[k], q[7], -TJ_ACCURACY, (double)W, D, zlo[k], zhi[k]);
#pragma unroll
            for (int r = 0; r < 4; ++r) rs[k][r] = (A)0;
        }
        const A wm1 = (A)(W - 1);
        for (int z0 = 0; z0 < D; z0 += zc) {
            const int zn = min(zc, D - z0);
            __syncthreads();
            if (tid < 2) {
                C4 zc4;
                zc4.x = zc4.y = zc4.z = zc4.w = (A)0;
                cell[zero + tid] = zc4;
            }
            // wave w stages slices w, w + NT/64, ...; lanes take consecutive
            // columns, SU columns per lane per batch so that 4*SU loads are in
            // flight before the first LDS write.  Rows outside the volume read
            // 0 (empty buffer range); columns past W read W-1.
            for (int zz = wave; zz < zn; zz += NT / 64) {
                const uint32_t zo = (uint32_t)((z0 + zz) * W);
                C4* dst = cell + zz * W1;
                for (int x0 = 0; x0 < W1; x0 += 64 * SU) {
                    A v[SU][4];
#pragma unroll
                    for (int u = 0; u < SU; ++u) {
                        const uint32_t o = zo + (uint32_t)min(x0 + 64 * u + lane, W - 1);
#pragma unroll
                        for (int r = 0; r < 4; ++r) v[u][r] = bld<A>(vol, rsc[r], o);
                    }
#pragma unroll
                    for (int u = 0; u < SU; ++u) {
                        const int xx = x0 + 64 * u + lane;
                        if (xx < W1) {
                            C4 c4;
                            c4.x = v[u][0];
                            c4.y = v[u][1];
                            c4.z = v[u][2];
                            c4.w = v[u][3];
                            dst[xx] = c4;
                        }
                    }
                }
            }
            __syncthreads();
            const int zlast = (z0 + zn == D) ? zn - 1 : zn;
            auto f64 = [&](auto ny, int z, int zo, A zl) {
                fp_slice_f64<A, G, decltype(ny)::value>(cell, z, zo, zero, zl, (double)wm1, ym, bd, td, zlo, zhi, rs);
            };
            if (neary)
                fp_slab<A, G, true>(cell, zn, zlast, z0, W1, zero, wm1, ym, base, tn, zlo, zhi, rs, f64);
            else
                fp_slab<A, G, false>(cell, zn, zlast, z0, W1, zero, wm1, ym, base, tn, zlo, zhi, rs, f64);
        }
        // projection, norm and difference per (projection, row); error partials
        double e2[G][4];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int p = p0 + k;
            const double* q = J.prm + PRM * min(p, n - 1);
            const double ninc = q[3];
            // the in-volume slices of a ray are consecutive: once it has entered,
            // it leaves (one more normIncrement) unless the last slice is inside
            const int cnt = zhi[k] - zlo[k] + 1;
            const double nrm = (cnt > 0) ? (double)(cnt + (zhi[k] == D - 1 ? 0 : 1)) * ninc : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + r;
                double v2 = 0.0;
                if (p < n && col && j < ph) {
                    const size_t o = ((size_t)p * ph + j) * pw + i;
                    const A norm = rin[r] ? (A)nrm : (A)0;
                    const A th = (norm > (A)0) ? (A)((double)rs[k][r] * ninc) : (A)NAN;
                    const A ex = ld<A>((const TE*)J.exps, o);
                    A d = (A)0;
                    if (!isnan(ex) && !isnan(th)) {
                        const double v = (double)((ex - th) / norm) * J.factor;
                        d = (A)v;
                        v2 = v * v;
                    }
                    st((A*)J.diff, o, d);
                }
                e2[k][r] = wave_sum(v2);
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int r = 0; r < 4; ++r) red[wave][k * 4 + r] = e2[k][r];
        }
        __syncthreads();
        if (tid < G * 4) {
            const int p = p0 + tid / 4, j = j0 + (tid & 3);
            if (p < n && j < ph) {
                double s = red[0][tid];
#pragma unroll
                for (int w = 1; w < NT / 64; ++w) s += red[w][tid];
                gp(J.err_part)[((size_t)p * ph + j) * J.nct + ct] = s;
            }
        }
    }
}

// Cells each k_job_fq thread stages per slab (held in registers while the
// previous slab is projected): with 4, at 4 rays per thread, they spill around
// the slice loop (a SIRT iteration at 512^3 7.31 against 6.78 ms with 3).
constexpr int FQ_CELLS = 3;

// The job forward projection with PG parts per block sharing a
// double-buffered slab (arithmetic in FP32): part h of a 1024-thread block
// follows the G rays of projections p0 + h G ... over the tile's 256 columns,
// and all parts project from one staged slab, so one staging serves PG * G
// projections.  Each thread owns the cells e = tid + q * (threads),
// q < FQ_CELLS, of every slab: their loads for slab k + 1 are issued before
// slab k is projected and stored into the other buffer after it (one barrier
// per slab).  The cells, the positions and the arithmetic are k_job_fp's.
template <typename T, typename TE, int G, int PG>
__global__ __launch_bounds__(NT * PG) void k_job_fq(const TjJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                                int ntiles, int lds_cells)
{
    using A = float;
    using C4 = float4;
    constexpr int BT = NT * PG;
    constexpr int QF = FQ_CELLS;
    extern __shared__ __align__(16) unsigned char smem[];
    C4* cells = reinterpret_cast<C4*>(smem);
    __shared__ double red[BT / 64][G * 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the part is uniform over a wave (NT is a multiple of 64): in a scalar
    // register the parameter rows of its projections are scalar loads
    const int part = __builtin_amdgcn_readfirstlane(tid / NT), ctid = tid - part * NT;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        if (tl.x < 0) continue;
        const TjJob& J = jobs[tl.x];
        const T* vol = (const T*)J.vol;
        const int W = J.W, D = J.D, pw = J.pw, ph = J.ph, n = J.n;
        const int W1 = W + 1;
        // slices per slab: QF cells per thread, two buffers of zc W1 + 2 cells
        const int zc = min(min(D, (QF * BT) / W1), (lds_cells / 2 - 2) / W1);
        const int bufc = zc * W1 + 2;
        const int zero = zc * W1;
        const int p0 = tl.y + part * G, j0 = tl.z * 4, ct = tl.w;
        const int i = ct * NT + ctid;
        const bool col = i < pw;
        const double ic = col ? i : pw - 1;
        const size_t plane = (size_t)D * W;
        bool rin[4];
        __amdgpu_buffer_rsrc_t rsc[4];
        A ym[4];
        bool neary = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int y = j0 + r + J.dy;
            rin[r] = (j0 + r < ph) && y >= 0 && y < J.H;
            rsc[r] = __builtin_amdgcn_make_buffer_rsrc((void*)(vol + (size_t)(rin[r] ? y : 0) * plane), 0,
                                                       rin[r] ? (int)(plane * sizeof(T)) : 0, 0x00020000);
            ym[r] = (y == J.H - 1) ? (A)0 : (A)1;
            neary = neary || (y == J.H - 1);
        }
        A base[G], tn[G], rs[G][4];
        int zlo[G], zhi[G];
        {
            double bd[G], td[G];
            ray_f64<G>(J.prm, p0, n, ic, bd, td);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                base[k] = (A)bd[k];
                tn[k] = (A)td[k];
                line_interval(bd[k], td[k], J.prm[PRM * min(p0 + k, n - 1) + 7], -TJ_ACCURACY, (double)W, D, zlo[k],
                              zhi[k]);
#pragma unroll
                for (int r = 0; r < 4; ++r) rs[k][r] = (A)0;
            }
        }
        // this thread's cells: slice zq of the slab (zc for none), element
        // offset oq within the slab's planes (columns past W read W - 1)
        int zq[QF];
        uint32_t oq[QF];
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int e = tid + q * BT;
            const int z = e / W1;
            zq[q] = min(z, zc);
            oq[q] = (uint32_t)(z * W + min(e - z * W1, W - 1));
        }
        A v[QF][4];
        auto fetch = [&](int zs) {
            const int zn = min(zc, D - zs);
            const uint32_t zo = (uint32_t)(zs * W);
#pragma unroll
            for (int q = 0; q < QF; ++q)
                if (zq[q] < zn) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[q][r] = bld<A>(vol, rsc[r], zo + oq[q]);
                }
        };
        auto store = [&](C4* c, int zs) {
            const int zn = min(zc, D - zs);
#pragma unroll
            for (int q = 0; q < QF; ++q)
                if (zq[q] < zn) c[tid + q * BT] = make_float4(v[q][0], v[q][1], v[q][2], v[q][3]);
        };
        fetch(0);
        if (tid < 4) cells[(tid >> 1) * bufc + zero + (tid & 1)] = make_float4(0.f, 0.f, 0.f, 0.f);
        store(cells, 0);
        __syncthreads();
        const A wm1 = (A)(W - 1);
        int buf = 0;
        for (int z0 = 0; z0 < D; z0 += zc) {
            const int zn = min(zc, D - z0);
            const bool more = z0 + zc < D;  // uniform
            if (more) fetch(z0 + zc);
            const int zlast = more ? zn : zn - 1;
            const C4* cur = cells + (size_t)buf * bufc;
            // the FP64 ray constants are read again where a slice needs them
            // (they would hold 4 G registers through the slab loop)
            auto f64 = [&](auto ny, int z, int zo, A zl) {
                double bd[G], td[G];
                ray_f64<G>(J.prm, p0, n, ic, bd, td);
                fp_slice_f64<A, G, decltype(ny)::value>(cur, z, zo, zero, zl, (double)wm1, ym, bd, td, zlo, zhi, rs);
            };
            if (neary)
                fp_slab<A, G, true, decltype(f64), true>(cur, zn, zlast, z0, W1, zero, wm1, ym, base, tn, zlo, zhi, rs,
                                                         f64);
            else
                fp_slab<A, G, false, decltype(f64), true>(cur, zn, zlast, z0, W1, zero, wm1, ym, base, tn, zlo, zhi,
                                                          rs, f64);
            if (more) {
                store(cells + (size_t)(buf ^ 1) * bufc, z0 + zc);
                __syncthreads();
                buf ^= 1;
            }
        }
        // projection, norm and difference per (projection, row); error
        // partials, each wave's sum stored as soon as it is formed
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int p = p0 + k;
            const double* q = J.prm + PRM * min(p, n - 1);
            const double ninc = q[3];
            const int cnt = zhi[k] - zlo[k] + 1;
            const double nrm = (cnt > 0) ? (double)(cnt + (zhi[k] == D - 1 ? 0 : 1)) * ninc : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + r;
                double v2 = 0.0;
                if (p < n && col && j < ph) {
                    const size_t o = ((size_t)p * ph + j) * pw + i;
                    const A norm = rin[r] ? (A)nrm : (A)0;
                    const A th = (norm > (A)0) ? (A)((double)rs[k][r] * ninc) : (A)NAN;
                    const A ex = ld<A>((const TE*)J.exps, o);
                    A d = (A)0;
                    if (!isnan(ex) && !isnan(th)) {
                        const double vv = (double)((ex - th) / norm) * J.factor;
                        d = (A)vv;
                        v2 = vv * vv;
                    }
                    st((A*)J.diff, o, d);
                }
                const double e2 = wave_sum(v2);
                if (lane == 0) red[wave][k * 4 + r] = e2;
            }
        }
        __syncthreads();
        // per part: the sum of its 4 waves in order
        if (tid < PG * G * 4) {
            const int h = tid / (G * 4), e = tid - h * (G * 4);
            const int p = tl.y + h * G + e / 4, j = j0 + (e & 3);
            if (p < n && j < ph) {
                double sum = red[4 * h][e];
#pragma unroll
                for (int w = 1; w < NT / 64; ++w) sum += red[4 * h + w][e];
                gp(J.err_part)[((size_t)p * ph + j) * J.nct + ct] = sum;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Forward projection for 16-bit storage on packed pairs.  Each LDS cell holds,
// for 8 volume rows, the two neighbours (v[x], v[x+1]) of a column as the two
// 16-bit halves of a word, so a row sample is one v_dot2 of that pair with the
// weights (1 - xf, xf) accumulated straight into the FP32 ray sum.  The ray
// sums stay FP32.
// ---------------------------------------------------------------------------
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef __bf16 b2_t __attribute__((ext_vector_type(2)));

template <typename T> struct Pair;
// Weights and pairs travel as 32-bit words and are reinterpreted only at the
// dot instruction: selects and constants of __bf16 vectors were miscompiled
// (a half of the vector dropped) where plain 32-bit values are not.
//
// The fraction is rounded to a multiple of 2^-11 (FP16) or 2^-8 (BF16): then
// both weights are exact in the 16-bit format and sum to exactly 1, so the
// sample's error is (q1 - q0) * (w1 - xf), bounded by the local difference
// times 2^-12 or 2^-9, and a sample of a constant region is exact.
// magic(W): adding it to a position in [0, W + 1] rounds the position to the
// weight grid (2^-11 for FP16, 2^-8 for BF16) in one FP32 addition, after which
// the integer part is the cell (plus the magic) and the fraction w1 is exact;
// of(w1) is the pair (1 - w1, w1).
template <> struct Pair<_Float16> {
    __device__ static __forceinline__ float magic(int W)
    {
        int m = 4096;
        while (m < W + 2) m <<= 1;
        return (float)m;
    }
    __device__ static __forceinline__ uint32_t of(float w1)
    {
        return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(1.0f - w1, w1));
    }
    __device__ static __forceinline__ uint32_t weights(float xf)
    {
        const float w1 = rintf(xf * 2048.0f) * (1.0f / 2048.0f);
        // both values are exact in FP16, so the truncating conversion is exact
        return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(1.0f - w1, w1));
    }
    __device__ static __forceinline__ uint32_t nearest() { return 0x00003C00u; }  // (1, 0)
    __device__ static __forceinline__ float dot(uint32_t p, uint32_t w, float acc)
    {
        return __builtin_amdgcn_fdot2(__builtin_bit_cast(h2_t, p), __builtin_bit_cast(h2_t, w), acc, false);
    }
};
template <> struct Pair<Bf16> {
    __device__ static __forceinline__ float magic(int) { return 32768.0f; }
    __device__ static __forceinline__ uint32_t of(float w1)
    {
        return __builtin_amdgcn_perm(__float_as_uint(w1), __float_as_uint(1.0f - w1), 0x07060302u);
    }
    __device__ static __forceinline__ uint32_t weights(float xf)
    {
        const float w1 = rintf(xf * 256.0f) * (1.0f / 256.0f);
        // exact in BF16: the high halves of the FP32 words are the BF16 values
        return __builtin_amdgcn_perm(__float_as_uint(w1), __float_as_uint(1.0f - w1), 0x07060302u);
    }
    // (1, 0).  Materialised in a register on purpose: as a literal the compiler
    // encodes it as the inline constant 1.0, which the BF16 dot reads as the
    // FP32 pattern 0x3F800000, i.e. the pair (0, 1).
    __device__ static __forceinline__ uint32_t nearest()
    {
        uint32_t v;
        asm volatile("v_mov_b32 %0, 0x3f80" : "=v"(v));
        return v;
    }
    __device__ static __forceinline__ float dot(uint32_t p, uint32_t w, float acc)
    {
        return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(b2_t, p), __builtin_bit_cast(b2_t, w), acc,
                                              false);
    }
};

template <typename T, int G>
__device__ __forceinline__ void fpp_accumulate(const uint4 (&c0)[G], const uint4 (&c1)[G], const uint32_t (&w)[G],
                                               float (&rs)[G][8])
{
#pragma unroll
    for (int k = 0; k < G; ++k) {
        rs[k][0] = Pair<T>::dot(c0[k].x, w[k], rs[k][0]);
        rs[k][1] = Pair<T>::dot(c0[k].y, w[k], rs[k][1]);
        rs[k][2] = Pair<T>::dot(c0[k].z, w[k], rs[k][2]);
        rs[k][3] = Pair<T>::dot(c0[k].w, w[k], rs[k][3]);
        rs[k][4] = Pair<T>::dot(c1[k].x, w[k], rs[k][4]);
        rs[k][5] = Pair<T>::dot(c1[k].y, w[k], rs[k][5]);
        rs[k][6] = Pair<T>::dot(c1[k].z, w[k], rs[k][6]);
        rs[k][7] = Pair<T>::dot(c1[k].w, w[k], rs[k][7]);
    }
}

// Slice z with FP64 positions (the nearest-value samples): `last` on the last
// slice (every row nearest), ymask bit r set for a last-row row r.  Cell c of
// the slice (offset zo) holds column x = c - 1; rays outside read cell 0 (zero).
// LDS index of cell c in a slice: even cells first, odd cells from HB on (HB
// is 4 mod 8 cells, so two neighbouring cells lie 64 bytes apart in the bank
// space and rays 1 or 2 cells apart read distinct banks)
__device__ __forceinline__ int fpp_cidx(int c, int hb) { return (c >> 1) + (c & 1) * hb; }

template <typename T, int G>
__device__ __forceinline__ void fpp_slice_f64(const uint4* cell, int hp, int hb, int z, int zo, bool last, double wd,
                                              unsigned ymask, const double* prm, int p0, int n, double ic,
                                              const int (&zlo)[G], const int (&zhi)[G], float (&rs)[G][8])
{
    int a[G];
    uint32_t w[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const double* q = prm + PRM * min(p0 + k, n - 1);
        const double xv = fma((double)z, q[2], ic * q[0] + q[1]);
        const double xs = fmin(fmax(xv, 0.0), wd);
        const int xb = (int)xs;
        w[k] = last ? Pair<T>::nearest() : Pair<T>::weights((float)(xs - (double)xb));
        a[k] = zo + fpp_cidx(((z >= zlo[k]) && (z <= zhi[k])) ? xb + 1 : 0, hb);
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const uint4 c0 = cell[a[k]], c1 = cell[hp + a[k]];
        const uint32_t v[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
        for (int r = 0; r < 8; ++r)
            rs[k][r] = Pair<T>::dot(v[r], ((ymask >> r) & 1u) ? Pair<T>::nearest() : w[k], rs[k][r]);
    }
}

// Projects one staged slab of the packed forward kernels: slices z0 .. z0 +
// zn - 1 of the tile's G rays.  cell: the slab's first cell (rows 0-3; rows
// 4-7 at + hp), ss cells per slice, odd cells from hb on.  Interpolated slices
// take FP32 positions rounded to the weight grid; the last volume slice (z0 +
// zn == D) takes FP64 positions with getPixel's nearest value.
template <typename T, int G>
__device__ __forceinline__ void fpp_project(const uint4* cell, int hp, int ss, int hb, int z0, int zn, int D, int W,
                                            unsigned ymask, const float (&base1)[G], const float (&tn)[G],
                                            const double* prm, int p0, int n, double ic, const int (&zlo)[G],
                                            const int (&zhi)[G], float (&rs)[G][8])
{
    const int zlast = (z0 + zn == D) ? zn - 1 : zn;
    const int zend = min(zlast, zn);
    const float wc1 = (float)(W + 1);
    // interpolated slices: FP32 positions; a last-row row (ymask) takes
    // getPixel's nearest value, the weights (1, 0)
    // LDS byte address of cell index 0 of the slab (uniform)
    const uint32_t sb = __builtin_amdgcn_readfirstlane((uint32_t)(size_t)(const TJ_LDS uint4*)cell);
    const float mg = Pair<T>::magic(W);
    const int mgh = (int)mg >> 1;
    auto slices = [&](auto nearest_rows) {
        for (int zz = 0; zz < zend; ++zz) {
            const float zf = (float)(z0 + zz);
            // byte address of index zz*ss - mg/2: cell c + mg then sits at
            // ((c + mg) >> 1) + ((c + mg) & 1) hb
            uint32_t sbz = sb + 16u * (uint32_t)(zz * ss - mgh);
            asm("" : "+s"(sbz));  // one v_lshl_add per ray below, not an add and a shift-add
            uint32_t a[G], w[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const float xs = __builtin_amdgcn_fmed3f(fmaf(zf, tn[k], base1[k]), 0.0f, wc1);
                const float xm = xs + mg;  // rounded to the weight grid
                const uint32_t cm = (uint32_t)xm;
                uint32_t odd = cm & 1u;
                asm("" : "+v"(odd));  // keeps (odd * hb) a 24-bit multiply-add, not a select
                uint32_t idx = __umul24(odd, (uint32_t)hb) + (cm >> 1);
                asm("" : "+v"(idx));
                a[k] = sbz + (idx << 4);
                w[k] = Pair<T>::of(__builtin_amdgcn_fractf(xm));
            }
            uint4 c0[G], c1[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                typedef uint32_t u4v __attribute__((ext_vector_type(4)));
                const u4v t0 = *(const TJ_LDS u4v*)(a[k]);
                const u4v t1 = *(const TJ_LDS u4v*)(a[k] + 16u * (uint32_t)hp);
                c0[k] = make_uint4(t0.x, t0.y, t0.z, t0.w);
                c1[k] = make_uint4(t1.x, t1.y, t1.z, t1.w);
            }
            if constexpr (decltype(nearest_rows)::value) {
                const uint32_t wn = Pair<T>::nearest();
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const uint32_t v[8] = {c0[k].x, c0[k].y, c0[k].z, c0[k].w,
                                           c1[k].x, c1[k].y, c1[k].z, c1[k].w};
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        rs[k][r] = Pair<T>::dot(v[r], ((ymask >> r) & 1u) ? wn : w[k], rs[k][r]);
                }
            } else {
                fpp_accumulate<T, G>(c0, c1, w, rs);
            }
        }
    };
    if (ymask) slices(std::true_type{});
    else slices(std::false_type{});
    if (zlast < zn)
        fpp_slice_f64<T, G>(cell, hp, hb, z0 + zlast, zlast * ss, true, (double)(W - 1), ymask, prm, p0, n, ic, zlo,
                            zhi, rs);
}

// Forward projection for 16-bit storage on packed pairs (BF16/FP16 forms).
// Tile (job, G projections, 8 detector rows, 256-column tile).  Per slice the
// LDS holds W + 2 cells of 32 bytes: cell c is column x = c - 1 of the 8 volume
// rows as pairs (v[x], v[x+1]) -- (v[W-1], v[W-1]) on the last column, getPixel's
// nearest value -- and cells 0 and W + 1 are zero.  An interpolated sample
// clamps its shifted position x + 1 to [0, W + 1] and reads the cell under it,
// so rays outside the volume read zeros and the inner loop has no compare:
//   xs = med3(fma(z, tan, base + 1), 0, W + 1), c = (int)xs, w = (1 - f, f)
// 10 VALU per ray and slice for 8 row samples, one v_dot2 each.  Inside/outside
// of a sample is decided by the FP32 position here (the ray norms still come
// from the exact FP64 interval); the last slice and the last volume row, where
// getPixel takes the nearest value, use FP64 positions as TomoJ computes them.
// Staging reads 32-bit words (two columns) when W is even, so a pair cell is
// a word or one byte permutation of two.
// CT: 256-column tiles per block (a block of CT * 256 threads stages the slab
// once for all of them)
template <typename T, typename TE, int G, int CT>
__global__ __launch_bounds__(NT * CT) void k_job_fpp(const TjJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                                int ntiles, int lds_cells)
{
    constexpr int RB = 8;  // volume rows per tile
    constexpr int PF = 3;  // staging items (column pairs of one slice) per thread and slab
    extern __shared__ __align__(16) unsigned char smem[];
    // cell[zz WC + c]: rows 0-3 of cell c of slice zz; cell[hp + zz WC + c]: rows 4-7
    // (two planes, so consecutive cells are consecutive 16-byte words)
    uint4* cell = reinterpret_cast<uint4*>(smem);
    constexpr int BT = NT * CT;  // threads per block
    __shared__ double red[BT / 64][G * RB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        if (tl.x < 0) continue;
        const TjJob& J = jobs[tl.x];
        const uint16_t* vol = (const uint16_t*)J.vol;
        const int W = J.W, D = J.D, pw = J.pw, ph = J.ph, n = J.n;
        const int WC = W + 2;                              // cells of a slice
        const int hb = (((WC + 1) / 2 + 3) & ~7) + 4;       // first odd cell: >= even count, 4 mod 8
        const int ss = hb + WC / 2;                        // slice stride (cells)
        const int zcl = min(D, lds_cells / (2 * ss));      // slices the LDS holds
        const int p0 = tl.y, j0 = tl.z * RB;
        const int i = tl.w * BT + tid;
        const bool col = i < pw;
        const double ic = col ? i : pw - 1;
        const size_t plane = (size_t)D * W;
        bool rin[RB];
        const uint16_t* rp[RB];
        unsigned ymask = 0;  // bit r: row r is the last volume row (nearest value in x)
        bool allin = true;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int y = j0 + r + J.dy;
            rin[r] = (j0 + r < ph) && y >= 0 && y < J.H;
            rp[r] = vol + (size_t)(rin[r] ? y : 0) * plane;
            ymask |= (y == J.H - 1) ? (1u << r) : 0u;
            allin = allin && rin[r];
        }
        const int half = W >> 1;
        // word staging (W even, every row inside, a slab within PF items per
        // thread) is software-pipelined: the next slab's words are loaded into
        // registers while the current slab is projected
        const bool words = allin && (W & 1) == 0 && half <= PF * BT;
        const int zc = words ? min(zcl, (PF * BT) / half) : zcl;
        const int hp = zc * ss;  // offset of the rows 4-7 plane
        float base1[G], tn[G], rs[G][RB];
        int zlo[G], zhi[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const double* q = J.prm + PRM * min(p0 + k, n - 1);
            const double bd = ic * q[0] + q[1];
            base1[k] = (float)(bd + 1.0);
            tn[k] = (float)q[2];
            line_interval(bd, q[2], q[7], -TJ_ACCURACY, (double)W, D, zlo[k], zhi[k]);
#pragma unroll
            for (int r = 0; r < RB; ++r) rs[k][r] = 0.0f;
        }
        // item q of this thread: m = tid + q BT = (slice zz, column pair x, x + 1)
        int izz[PF], ix[PF];
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int m = tid + q * BT;
            izz[q] = words ? m / half : 0;
            ix[q] = words ? 2 * (m - izz[q] * half) : 0;
        }
        // the volume rows as buffers (uniform, so the loads take a scalar base
        // and a 32-bit offset)
        __amdgpu_buffer_rsrc_t rsc[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r)
            rsc[r] = __builtin_amdgcn_make_buffer_rsrc((void*)rp[r], 0, (int)(plane * 2), 0x00020000);
        uint32_t wa[PF][RB], wb[PF][RB];
        // cell x + 1 is the word at x; cell x + 2 is (v[x+1], v[x+2]) from the
        // words at x and x + 2, or (v[W-1], v[W-1]) on the last pair
        auto fetch = [&](int zs, int znn) {
#pragma unroll
            for (int q = 0; q < PF; ++q) {
                const bool ok = izz[q] < znn;
                const uint32_t o0 = (uint32_t)((zs + (ok ? izz[q] : 0)) * W + ix[q]) * 2u;
                const uint32_t o1 = (ix[q] + 2 >= W) ? o0 : o0 + 4u;
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    wa[q][r] = __builtin_amdgcn_raw_buffer_load_b32(rsc[r], o0, 0, 0);
                    wb[q][r] = __builtin_amdgcn_raw_buffer_load_b32(rsc[r], o1, 0, 0);
                }
            }
        };
        if (words) fetch(0, min(zc, D));
        for (int z0 = 0; z0 < D; z0 += zc) {
            const int zn = min(zc, D - z0);
            __syncthreads();
            for (int q = tid; q < zn; q += BT) {
                uint4* d = cell + (size_t)q * ss;
                const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
                d[fpp_cidx(0, hb)] = zero4;
                d[hp + fpp_cidx(0, hb)] = zero4;
                d[fpp_cidx(W + 1, hb)] = zero4;
                d[hp + fpp_cidx(W + 1, hb)] = zero4;
            }
            if (words) {
#pragma unroll
                for (int q = 0; q < PF; ++q) {
                    if (izz[q] < zn) {
                        const uint32_t sel = (ix[q] + 2 >= W) ? 0x03020302u : 0x05040302u;
                        // cell x + 1 (odd) and cell x + 2 (even)
                        uint4* d1 = cell + ((size_t)izz[q] * ss + hb + (ix[q] >> 1));
                        uint4* d2 = cell + ((size_t)izz[q] * ss + (ix[q] >> 1) + 1);
                        d1[0] = make_uint4(wa[q][0], wa[q][1], wa[q][2], wa[q][3]);
                        d1[hp] = make_uint4(wa[q][4], wa[q][5], wa[q][6], wa[q][7]);
                        d2[0] = make_uint4(__builtin_amdgcn_perm(wb[q][0], wa[q][0], sel),
                                           __builtin_amdgcn_perm(wb[q][1], wa[q][1], sel),
                                           __builtin_amdgcn_perm(wb[q][2], wa[q][2], sel),
                                           __builtin_amdgcn_perm(wb[q][3], wa[q][3], sel));
                        d2[hp] = make_uint4(__builtin_amdgcn_perm(wb[q][4], wa[q][4], sel),
                                            __builtin_amdgcn_perm(wb[q][5], wa[q][5], sel),
                                            __builtin_amdgcn_perm(wb[q][6], wa[q][6], sel),
                                            __builtin_amdgcn_perm(wb[q][7], wa[q][7], sel));
                    }
                }
                if (z0 + zn < D) fetch(z0 + zn, min(zc, D - z0 - zn));
            } else {
                // any width, rows outside the volume (they read row 0 and store 0)
                for (int m = tid; m < zn * W; m += BT) {
                    const int zz = m / W, x = m - zz * W;
                    const size_t oa = (size_t)(z0 + zz) * W + x, ob = (size_t)(z0 + zz) * W + min(x + 1, W - 1);
                    uint32_t v[RB];
#pragma unroll
                    for (int r = 0; r < RB; ++r)
                        v[r] = rin[r] ? ((uint32_t)gp(rp[r])[oa] | ((uint32_t)gp(rp[r])[ob] << 16)) : 0u;
                    uint4* d = cell + ((size_t)zz * ss + fpp_cidx(x + 1, hb));
                    d[0] = make_uint4(v[0], v[1], v[2], v[3]);
                    d[hp] = make_uint4(v[4], v[5], v[6], v[7]);
                }
            }
            __syncthreads();
            fpp_project<T, G>(cell, hp, ss, hb, z0, zn, D, W, ymask, base1, tn, J.prm, p0, n, ic, zlo, zhi, rs);
        }
        // projection, norm and difference per (projection, row); error partials
        double e2[G][RB];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int p = p0 + k;
            const double ninc = J.prm[PRM * min(p, n - 1) + 3];
            const int cnt = zhi[k] - zlo[k] + 1;
            const double nrm = (cnt > 0) ? (double)(cnt + (zhi[k] == D - 1 ? 0 : 1)) * ninc : 0.0;
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int j = j0 + r;
                double v2 = 0.0;
                if (p < n && col && j < ph) {
                    const size_t o = ((size_t)p * ph + j) * pw + i;
                    const float norm = rin[r] ? (float)nrm : 0.0f;
                    const float th = (norm > 0.0f) ? (float)((double)rs[k][r] * ninc) : NAN;
                    const float ex = ld<float>((const TE*)J.exps, o);
                    float d = 0.0f;
                    if (!isnan(ex) && !isnan(th)) {
                        const double v = (double)((ex - th) / norm) * J.factor;
                        d = (float)v;
                        v2 = v * v;
                    }
                    st((float*)J.diff, o, d);
                }
                e2[k][r] = wave_sum(v2);
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int r = 0; r < RB; ++r) red[wave][k * RB + r] = e2[k][r];
        }
        __syncthreads();
        // per 256-column tile: the sum of its 4 waves in order
        if (tid < CT * G * RB) {
            const int h = tid / (G * RB), e = tid - h * (G * RB);
            const int p = p0 + e / RB, j = j0 + (e % RB), cth = tl.w * CT + h;
            if (p < n && j < ph && cth < J.nct) {
                double s = red[4 * h][e];
#pragma unroll
                for (int w = 1; w < NT / 64; ++w) s += red[4 * h + w][e];
                gp(J.err_part)[((size_t)p * ph + j) * J.nct + cth] = s;
            }
        }
    }
}

// Cells per plane of a k_job_fpq slab buffer.  A compile-time plane stride
// lets the read of a cell's rows 4-7 take the stride as the LDS instruction's
// immediate offset (16 x 2432 bytes); the host gives the kernel the 4 planes of
// its two buffers (152 KiB).
constexpr int FPQ_PLANE = 2432;

// The packed forward projection with a double-buffered slab.  A block of
// PG * 256 threads holds PG parts, each the G rays of its own projections
// (p0 + part G ...) over the tile's 256 columns, and all parts share the
// staged slab, so one staging serves PG * G * 256 rays.  The slab alternates
// between two LDS buffers: the next slab's words are loaded into registers
// before the current slab is projected and written into the other buffer
// after it (one barrier per slab).  Each thread stages one column pair of one
// slice per slab (zc = threads / (W / 2) slices, at most what the LDS holds);
// rows outside the volume or the detector read zeros through the buffer range
// check.  Requires an even W; the cells and the arithmetic are k_job_fpp's.
// The slab planes are FPQ_PLANE cells apart (zc ss <= FPQ_PLANE).
template <typename T, typename TE, int G, int PG>
__global__ __launch_bounds__(NT * PG) void k_job_fpq(const TjJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                                int ntiles, int /*lds_cells: 4 FPQ_PLANE*/)
{
    constexpr int RB = 8;
    constexpr int BT = NT * PG;
    extern __shared__ __align__(16) unsigned char smem[];
    uint4* cells = reinterpret_cast<uint4*>(smem);
    __shared__ double red[BT / 64][G * RB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int part = tid / NT, ctid = tid - part * NT;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        if (tl.x < 0) continue;
        const TjJob& J = jobs[tl.x];
        const uint16_t* vol = (const uint16_t*)J.vol;
        const int W = J.W, D = J.D, pw = J.pw, ph = J.ph, n = J.n;
        const int WC = W + 2;
        const int hb = (((WC + 1) / 2 + 3) & ~7) + 4;
        const int ss = hb + WC / 2;
        const int half = W >> 1;
        // slices per slab: one column pair per thread, as many as a plane holds
        const int zc = min(min(D, BT / half), FPQ_PLANE / ss);
        constexpr int hp = FPQ_PLANE;    // rows 4-7 plane within a buffer
        constexpr int bufc = 2 * hp;     // cells per buffer
        const int p0 = tl.y + part * G, j0 = tl.z * RB;
        const int i = tl.w * NT + ctid;
        const bool col = i < pw;
        const double ic = col ? i : pw - 1;
        const size_t plane = (size_t)D * W;
        unsigned ymask = 0;
        unsigned rmask = 0;  // bit r: row r is inside the volume and the detector
        __amdgpu_buffer_rsrc_t rsc[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int y = j0 + r + J.dy;
            const bool in = (j0 + r < ph) && y >= 0 && y < J.H;
            rmask |= in ? (1u << r) : 0u;
            ymask |= (y == J.H - 1) ? (1u << r) : 0u;
            rsc[r] = __builtin_amdgcn_make_buffer_rsrc((void*)(vol + (size_t)(in ? y : 0) * plane), 0,
                                                       in ? (int)(plane * 2) : 0, 0x00020000);
        }
        float base1[G], tn[G], rs[G][RB];
        int zlo[G], zhi[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const double* q = J.prm + PRM * min(p0 + k, n - 1);
            const double bd = ic * q[0] + q[1];
            base1[k] = (float)(bd + 1.0);
            tn[k] = (float)q[2];
            line_interval(bd, q[2], q[7], -TJ_ACCURACY, (double)W, D, zlo[k], zhi[k]);
#pragma unroll
            for (int r = 0; r < RB; ++r) rs[k][r] = 0.0f;
        }
        // this thread's item: slice izz of the slab, column pair (ix, ix + 1)
        const int izz = tid / half;
        const int ix = 2 * (tid - izz * half);
        const uint32_t sel = (ix + 2 >= W) ? 0x03020302u : 0x05040302u;
        uint32_t wa[RB], wb[RB];
        // cell x + 1 is the word at x; cell x + 2 is (v[x+1], v[x+2]) from the
        // words at x and x + 2, or (v[W-1], v[W-1]) on the last pair
        auto fetch = [&](int zs) {
            if (izz < min(zc, D - zs)) {
                const uint32_t o0 = (uint32_t)((zs + izz) * W + ix) * 2u;
                const uint32_t o1 = (ix + 2 >= W) ? o0 : o0 + 4u;
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    wa[r] = __builtin_amdgcn_raw_buffer_load_b32(rsc[r], o0, 0, 0);
                    wb[r] = __builtin_amdgcn_raw_buffer_load_b32(rsc[r], o1, 0, 0);
                }
            }
        };
        auto store = [&](uint4* c, int zs) {
            if (izz < min(zc, D - zs)) {
                uint4* d1 = c + ((size_t)izz * ss + hb + (ix >> 1));
                uint4* d2 = c + ((size_t)izz * ss + (ix >> 1) + 1);
                d1[0] = make_uint4(wa[0], wa[1], wa[2], wa[3]);
                d1[hp] = make_uint4(wa[4], wa[5], wa[6], wa[7]);
                d2[0] = make_uint4(__builtin_amdgcn_perm(wb[0], wa[0], sel), __builtin_amdgcn_perm(wb[1], wa[1], sel),
                                   __builtin_amdgcn_perm(wb[2], wa[2], sel), __builtin_amdgcn_perm(wb[3], wa[3], sel));
                d2[hp] = make_uint4(__builtin_amdgcn_perm(wb[4], wa[4], sel), __builtin_amdgcn_perm(wb[5], wa[5], sel),
                                    __builtin_amdgcn_perm(wb[6], wa[6], sel), __builtin_amdgcn_perm(wb[7], wa[7], sel));
            }
        };
        fetch(0);
        // the zero cells 0 and W + 1 of every slice of both buffers (the
        // previous tile's readers passed its final barrier)
        for (int q = tid; q < 2 * zc; q += BT) {
            uint4* d = cells + (size_t)(q / zc) * bufc + (size_t)(q % zc) * ss;
            const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
            d[fpp_cidx(0, hb)] = zero4;
            d[hp + fpp_cidx(0, hb)] = zero4;
            d[fpp_cidx(W + 1, hb)] = zero4;
            d[hp + fpp_cidx(W + 1, hb)] = zero4;
        }
        store(cells, 0);
        __syncthreads();
        int buf = 0;
        for (int z0 = 0; z0 < D; z0 += zc) {
            const int zn = min(zc, D - z0);
            const bool more = z0 + zc < D;  // uniform
            if (more) fetch(z0 + zc);
            fpp_project<T, G>(cells + (size_t)buf * bufc, hp, ss, hb, z0, zn, D, W, ymask, base1, tn, J.prm, p0, n,
                              ic, zlo, zhi, rs);
            if (more) {
                store(cells + (size_t)(buf ^ 1) * bufc, z0 + zc);
                __syncthreads();
                buf ^= 1;
            }
        }
        // projection, norm and difference per (projection, row); error partials
        double e2[G][RB];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int p = p0 + k;
            const double ninc = J.prm[PRM * min(p, n - 1) + 3];
            const int cnt = zhi[k] - zlo[k] + 1;
            const double nrm = (cnt > 0) ? (double)(cnt + (zhi[k] == D - 1 ? 0 : 1)) * ninc : 0.0;
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int j = j0 + r;
                double v2 = 0.0;
                if (p < n && col && j < ph) {
                    const size_t o = ((size_t)p * ph + j) * pw + i;
                    const float norm = ((rmask >> r) & 1u) ? (float)nrm : 0.0f;
                    const float th = (norm > 0.0f) ? (float)((double)rs[k][r] * ninc) : NAN;
                    const float ex = ld<float>((const TE*)J.exps, o);
                    float d = 0.0f;
                    if (!isnan(ex) && !isnan(th)) {
                        const double v = (double)((ex - th) / norm) * J.factor;
                        d = (float)v;
                        v2 = v * v;
                    }
                    st((float*)J.diff, o, d);
                }
                e2[k][r] = wave_sum(v2);
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int r = 0; r < RB; ++r) red[wave][k * RB + r] = e2[k][r];
        }
        __syncthreads();
        // per part: the sum of its 4 waves in order
        if (tid < PG * G * RB) {
            const int h = tid / (G * RB), e = tid - h * (G * RB);
            const int p = tl.y + h * G + e / RB, j = j0 + (e % RB);
            if (p < n && j < ph && tl.w < J.nct) {
                double sum = red[4 * h][e];
#pragma unroll
                for (int w = 1; w < NT / 64; ++w) sum += red[4 * h + w][e];
                gp(J.err_part)[((size_t)p * ph + j) * J.nct + tl.w] = sum;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Backprojection of all N differences of a job (SIRT), added to each voxel in
// the job's projection order (orderShuffleART), as osartIteration adds them.  Tile (job, 256-column
// tile, R volume rows, KC slices): each thread keeps the KC*R sums of its voxel
// column in registers; per group of G images the rows those voxels reach are
// staged in LDS as pairs (ll, lr - ll) over the reachable detector columns.
// ---------------------------------------------------------------------------
template <typename T, typename A, int R, int KC, int G, int MODE>
__global__ __launch_bounds__(NT) void k_job_bp(const TjJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                               int ntiles)
{
    using C2 = typename V2<A>::t;
    constexpr bool POS = (MODE & 1) != 0;     // positivity after each image
    constexpr bool MASTER = (MODE & 2) != 0;  // FP32 master volume beside the stored one
    // FP32 cell positions (16-bit storage forms): LDS cell li of an image row
    // holds column x = x0 + li - 1, zero outside [0, pw); a voxel clamps its
    // shifted position xx + 1 to [0, pw + 1] and reads the cell under it, so
    // the slice loop has no compare (inside/outside decided in FP32).  The
    // cells are the FP16 pairs (d[x], d[x+1] - d[x]) of the job's pair image,
    // so an update is one v_dot2 with the weights (1, f) into the FP32 sum, and
    // the next group's cells are loaded into registers while the current group
    // is backprojected.
    constexpr bool FAST = (MODE & 4) != 0;
    constexpr int XW = NT + KC + 8;  // reachable columns + a zero cell at XW-1
    static_assert(KC % 4 == 0, "slices per tile come in fours");
    constexpr int RPW = (G * R + 3) / 4;  // staged rows per wave
    constexpr int NB = (XW + 63) / 64;    // 64-cell batches per row
    constexpr int XWP = NB * 64;          // FP16-pair row stride (whole batches)
    // FP32/FP64 pairs [G][R][XW]; or, for the FP16-pair form, two buffers of
    // [G][R][XWP] words filled by direct buffer-to-LDS loads
    constexpr size_t LDS_BYTES = FAST ? (size_t)2 * G * R * XWP * 4 : (size_t)G * R * XW * sizeof(C2);
    __shared__ __align__(16) unsigned char lraw[LDS_BYTES];
    C2(*lds)[R][XW] = reinterpret_cast<C2(*)[R][XW]>(lraw);
    uint32_t(*ldsf)[G][R][XWP] = reinterpret_cast<uint32_t(*)[G][R][XWP]>(lraw);
    __shared__ double s_c[2][G], s_ms[2][G], s_b2[2][G], s_ims[2][G];
    __shared__ int s_x0[2][G], s_img[2][G];
    __shared__ int s_row[2][G][R];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        if (tl.x < 0) continue;
        const TjJob& J = jobs[tl.x];
        T* vol = (T*)J.vol;
        const A* diff = (const A*)J.diff;
        const int W = J.W, D = J.D, pw = J.pw, ph = J.ph, n = J.n;
        const int i0 = tl.y * NT, i = i0 + tid;
        const bool active = i < W;
        const int ilast = min(i0 + NT - 1, W - 1);
        const double ie = active ? i : ilast;
        const int j0 = tl.z * R, nr = min(R, J.H - j0);
        const int k0 = tl.w * KC, nk = min(KC, D - k0);
        const size_t plane = (size_t)D * W;
        A acc[KC][R];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk)
#pragma unroll
            for (int r = 0; r < R; ++r)
                acc[kk][r] = (active && kk < nk && r < nr)
                                 ? (MASTER ? ld<A>(J.master, (size_t)(j0 + r) * plane + (size_t)(k0 + kk) * W + i)
                                           : ld<A>(vol, (size_t)(j0 + r) * plane + (size_t)(k0 + kk) * W + i))
                                 : (A)0;
        // constants of the images p0 .. p0 + gn - 1 into parameter buffer `buf`
        auto params = [&](int p0, int gn, int buf) {
            if (tid < gn) {
                const int pj = J.order[p0 + tid];
                s_img[buf][tid] = pj;
                const double* q = J.prm + PRM * pj;
                const double c = q[4], ms = q[5], b2 = q[6];
                // smallest xx of the tile (xx is linear in i and k: a corner);
                // one column of margin for the rounding of the arithmetic type
                const double ka = k0, kb = k0 + KC - 1;
                const double m = fmin(fmin((double)i0 * c + ka * ms, (double)i0 * c + kb * ms),
                                      fmin((double)ilast * c + ka * ms, (double)ilast * c + kb * ms)) + b2;
                s_x0[buf][tid] = FAST ? min(max(0, (int)floor(m)), pw + 1) : max(0, (int)floor(m) - 1);
                s_c[buf][tid] = c;
                s_ms[buf][tid] = ms;
                s_b2[buf][tid] = b2;
                s_ims[buf][tid] = q[8];
            }
            if (tid < G * R) {
                const int g = tid / R, r = tid % R;
                const int jj = j0 + r + J.dyb;
                s_row[buf][g][r] = (g < gn && r < nr && jj >= 0 && jj < ph) ? jj : -1;
            }
        };
        if constexpr (FAST) {
            // wave w stages rows gr = w, w + 4, ... of the group: batch b of a
            // row is cells 64 b + lane, loaded from the pair image straight into
            // LDS buffer `buf` by range-checked buffer loads (cells off the row,
            // or rows outside the detector, read 0)
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            auto fetch = [&](int buf, int gn) {
#pragma unroll
                for (int w = 0; w < RPW; ++w) {
                    const int gr = wv + 4 * w;
                    if (gr < G * R) {
                        const int g = gr / R, r = gr % R;
                        const int jj = __builtin_amdgcn_readfirstlane(s_row[buf][g][r]);
                        const int xsh = __builtin_amdgcn_readfirstlane(s_x0[buf][g]) - 1;
                        const int img = __builtin_amdgcn_readfirstlane(s_img[buf][g]);
                        const bool okr = gr < gn * R && jj >= 0;
                        const uint32_t* prow = J.pairs + ((size_t)(okr ? img : 0) * ph + (okr ? jj : 0)) * pw;
                        const __amdgpu_buffer_rsrc_t rs =
                            __builtin_amdgcn_make_buffer_rsrc((void*)prow, 0, okr ? pw * 4 : 0, 0x00020000);
#pragma unroll
                        for (int bb = 0; bb < NB; ++bb)
                            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                                rs, (TJ_LDS void*)&ldsf[buf][g][r][64 * bb], 4,
                                (uint32_t)(xsh + 64 * bb + lane) * 4u, 0, 0, 0);
                    }
                }
            };
            __syncthreads();  // the previous tile is done with the parameters and cells
            params(0, min(G, n), 0);
            __syncthreads();
            fetch(0, min(G, n));
            int buf = 0;
            for (int p0 = 0; p0 < n; p0 += G, buf ^= 1) {
                const int gn = min(G, n - p0);
                __builtin_amdgcn_s_waitcnt(0);  // this thread's loads into buffer `buf` are done
                __syncthreads();  // all of them; the other buffer and parameters are free
                if (p0 + G < n) params(p0 + G, min(G, n - p0 - G), buf ^ 1);
                __syncthreads();
                if (p0 + G < n) fetch(buf ^ 1, min(G, n - p0 - G));
                for (int g = 0; g < gn; ++g) {
                    const float ms = (float)s_ms[buf][g];
                    const float xe1 = (float)(ie * s_c[buf][g] + s_b2[buf][g] + 1.0);
                    const float cmax = (float)(pw + 1);
                    // byte address of shifted cell 0 of row 0 (LDS addresses are
                    // 32-bit; uniform, so it lives in a scalar register and a cell
                    // address is one shift-add)
                    const uint32_t cg = __builtin_amdgcn_readfirstlane(
                        (uint32_t)(size_t)(const TJ_LDS uint32_t*)&ldsf[buf][g][0][0] - 4u * (uint32_t)s_x0[buf][g]);
                    // 1.0 in a scalar register: with both halves converted from FP32
                    // registers the weight pair is one packed conversion
                    float one = 1.0f;
                    asm("" : "+s"(one));
                    // 4 slices at a time: positions, then all LDS reads, then the sums
#pragma unroll
                    for (int k4 = 0; k4 < KC; k4 += 4) {
                        int c[4];
                        h2_t wt[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float xs =
                                __builtin_amdgcn_fmed3f(fmaf((float)(k0 + k4 + u), ms, xe1), 0.0f, cmax);
                            c[u] = (int)xs;
                            wt[u] = h2_t{(_Float16)one, (_Float16)__builtin_amdgcn_fractf(xs)};
                        }
                        uint32_t q[4][R];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int r = 0; r < R; ++r)
                                q[u][r] = ((const TJ_LDS uint32_t*)(cg + 4u * (uint32_t)c[u]))[r * XWP];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                A a = __builtin_amdgcn_fdot2(__builtin_bit_cast(h2_t, q[u][r]), wt[u],
                                                             acc[k4 + u][r], false);
                                if (POS) a = (a < (A)0) ? (A)0 : a;
                                acc[k4 + u][r] = a;
                            }
                    }
                }
            }
        } else {
            const A pwm1 = (A)(pw - 1);
            for (int p0 = 0; p0 < n; p0 += G) {
                const int gn = min(G, n - p0);
                __syncthreads();
                params(p0, gn, 0);
                __syncthreads();
                for (int gr = wave; gr < gn * R; gr += NT / 64) {
                    const int g = gr / R, r = gr - (gr / R) * R;
                    const int jj = s_row[0][g][r], x0 = s_x0[0][g];
                    const A* row = diff + ((size_t)s_img[0][g] * ph + max(jj, 0)) * pw;
                    C2* dst = &lds[g][r][0];
                    // SU columns per lane per batch, loads first (addresses clamped
                    // into the row; cells that take no value store 0)
                    for (int c0 = 0; c0 < XW; c0 += 64 * SU) {
                        A ll[SU], lr[SU];
#pragma unroll
                        for (int u = 0; u < SU; ++u) {
                            const int xs = x0 + c0 + 64 * u + lane;
                            ll[u] = ld<A>(row, min(xs, pw - 1));
                            lr[u] = ld<A>(row, min(xs + 1, pw - 1));
                        }
#pragma unroll
                        for (int u = 0; u < SU; ++u) {
                            const int xx = c0 + 64 * u + lane;
                            if (xx < XW) {
                                const bool take = jj >= 0 && x0 + xx < pw && xx < XW - 1;
                                C2 cv;
                                cv.x = take ? ll[u] : (A)0;
                                cv.y = take ? lr[u] - ll[u] : (A)0;
                                dst[xx] = cv;
                            }
                        }
                    }
                }
                __syncthreads();
                for (int g = 0; g < gn; ++g) {
                    const A ms = (A)s_ms[0][g];
                    const double xed = ie * s_c[0][g] + s_b2[0][g];
                    const A xe = (A)xed;
                    const int x0 = s_x0[0][g];
                    // slices whose xx lies in [0, pw) (VoxelProjector3D.java:285)
                    int klo, khi;
                    line_interval(xed, s_ms[0][g], s_ims[0][g], 0.0, (double)pw, D, klo, khi);
#pragma unroll
                    for (int kk = 0; kk < KC; ++kk) {
                        const A xx = fma((A)(k0 + kk), ms, xe);
                        const bool vx = (k0 + kk >= klo) && (k0 + kk <= khi);
                        const A xs = vx ? fmin(fmax(xx, (A)0), pwm1) : (A)0;
                        const int xb = (int)xs;
                        const A xf = xs - (A)xb;
                        const int li = vx ? xb - x0 : XW - 1;
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const C2 q = lds[g][r][li];
                            A a = acc[kk][r] + fma(xf, q.y, q.x);
                            if (POS) a = (a < (A)0) ? (A)0 : a;
                            acc[kk][r] = a;
                        }
                    }
                }
            }
        }
        if (active) {
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (kk < nk && r < nr) {
                        A a = acc[kk][r];
                        if (isnan(a) || isinf(a)) a = (A)0;
                        const size_t o = (size_t)(j0 + r) * plane + (size_t)(k0 + kk) * W + i;
                        if (MASTER) st(J.master, o, (float)a);
                        st(vol, o, a);
                    }
        }
    }
}

// Sum of squared differences of each job of a run, in a fixed order; appended
// to the job's error history.  One block per job.
constexpr int NTE = 1024;  // threads of the error reduction (one block per job)
__global__ __launch_bounds__(NTE) void k_job_err(const TjJob* __restrict__ jobs, int njobs)
{
    __shared__ double sh[NTE / 64];
    const TjJob& J = jobs[blockIdx.x];
    const size_t m = (size_t)J.n * J.ph * J.nct;
    // 8 independent partial sums per thread, so 8 loads are in flight
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t k = threadIdx.x;
    for (; k + 7 * NTE < m; k += 8 * NTE) {
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] += gp(J.err_part)[k + (size_t)u * NTE];
    }
    for (int u = 0; k < m; k += NTE, ++u) p[u] += gp(J.err_part)[k];
    double s = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = sh[0];
        for (int w = 1; w < NTE / 64; ++w) tot += sh[w];
        const int it = *J.iter;
        J.err_hist[it % J.hist_cap] = tot;
        *J.iter = it + 1;
    }
}

// Differences of the 16-bit forms as FP16 pairs (d[x], d[x+1] - d[x]) for the
// backprojection's cells; (d[pw-1], 0) on the last column (ImageJ's clamped
// edge).  Rounded to nearest; magnitudes beyond the FP16 range saturate.
// Blocks walk detector rows (all images) of job blockIdx.y.
__global__ __launch_bounds__(NT) void k_job_pairs(const TjJob* __restrict__ jobs)
{
    const TjJob& J = jobs[blockIdx.y];
    const float* diff = (const float*)J.diff;
    const int pw = J.pw;
    const size_t rows = (size_t)J.n * J.ph;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const size_t o = row * pw;
        for (int x = threadIdx.x; x < pw; x += NT) {
            const float d0 = gp(diff)[o + x];
            const float d1 = (x + 1 < pw) ? gp(diff)[o + x + 1] : d0;
            const float lim = 65504.0f;
            const h2_t h = {(_Float16)fminf(fmaxf(d0, -lim), lim), (_Float16)fminf(fmaxf(d1 - d0, -lim), lim)};
            gp(J.pairs)[o + x] = __builtin_bit_cast(uint32_t, h);
        }
    }
}

// Storage conversions and layout changes.
template <typename TI, typename TO>
__global__ void k_convert(const TI* __restrict__ src, TO* __restrict__ dst, size_t n)
{
    using A = typename std::conditional<std::is_same<TO, double>::value, double, float>::type;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        st(dst, k, ld<A>(src, k));
}

// host [z][y][x] float -> device [y][z][x] T, and back
template <typename T>
__global__ void k_zyx_to_job(const float* __restrict__ src, T* __restrict__ vol, int W, int H, int D)
{
    using A = typename std::conditional<std::is_same<T, double>::value, double, float>::type;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= W) return;
    st(vol, ((size_t)y * D + z) * W + x, (A)src[((size_t)z * H + y) * W + x]);
}
template <typename T>
__global__ void k_job_to_zyx(const T* __restrict__ vol, float* __restrict__ dst, int W, int H, int D)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= W) return;
    dst[((size_t)z * H + y) * W + x] = ld<float>(vol, ((size_t)y * D + z) * W + x);
}

}  // namespace tjb

// ===========================================================================
// host side
// ===========================================================================
using tjb::Bf16;
using tjb::TjJob;

static int bfail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    tj_record_error(buf);
    return code;
}

#define TB_HIP(call)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return bfail(TJ_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                           __FILE__, __LINE__);                                     \
    } while (0)

static size_t storage_bytes(int s)
{
    switch (s) {
    case TJ_STORAGE_F32: return 4;
    case TJ_STORAGE_BF16: return 2;
    case TJ_STORAGE_F16: return 2;
    case TJ_STORAGE_F64: return 8;
    default: return 0;
    }
}
static size_t arith_bytes(int s) { return s == TJ_STORAGE_F64 ? 8 : 4; }
static bool half_storage(int s) { return s == TJ_STORAGE_BF16 || s == TJ_STORAGE_F16; }
// storage type of a job's measured projections
static int proj_storage(const tj_job_desc& d)
{
    return (half_storage(d.storage) && d.fp32_projections) ? TJ_STORAGE_F32 : d.storage;
}
static int fp_col_tiles(int pw) { return (pw + tjb::NT - 1) / tjb::NT; }

struct BJob {
    int id = 0;
    tj_job_desc d{};
    void* vol = nullptr;
    float* master = nullptr;  // FP32 master volume (16-bit storage with fp32_master)
    uint32_t* pairs = nullptr;  // FP16 difference pairs (16-bit storage)
    void* exps = nullptr;
    void* diff = nullptr;
    double* err_part = nullptr;
    double* prm = nullptr;
    int* order = nullptr;
    double* err_hist = nullptr;
    int* iter = nullptr;
    std::vector<double> euler;  // n x 9
    std::vector<double> hprm;   // n x PRM
    int dy = 0, dyb = 0;
    size_t bytes = 0;
};

// A launch resolved, from the storage and the tuning, to one kernel instance
// with its block size and dynamic LDS.
struct Kern {
    const void* fn = nullptr;
    int block = tjb::NT;
    int lds = 0;     // dynamic LDS bytes
    int cells = -1;  // forward kernels: LDS cells (the kernel's last argument); -1 for the backprojection
};

struct RunPlan {
    std::vector<int> ids;
    int iterations = 0;
    tj_batch_tuning tune{};
    TjJob* djobs = nullptr;
    std::vector<int4*> dtiles;  // owned tile tables
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    // launches of one iteration: kind 0 forward, 1 error, 3 difference pairs, 2 back
    struct L {
        int kind, storage, pos, ntiles, grid;  // pos: back mode bits (back), widest volume (forward)
        int4* tiles;
        int job0, njobs;  // error launch: jobs [job0, job0+njobs) of djobs
        int pstorage;     // forward launch: storage of the measured projections
        Kern k;           // forward and back launches: the kernel
    };
    std::vector<L> launches;
};

struct tj_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0, used = 0;
    int next_id = 1;
    std::map<int, BJob*> jobs;
    // the autotuner's pick at 512^3 BF16; the FP32 forward on the shared
    // double-buffered slab (k_job_fq: per SIRT iteration 6.78 against 7.28 ms
    // with k_job_fp at 512^3, 31.4 against 43.7 ms at 1024x1024x512)
    tj_batch_tuning tune{4, 128, 4, 16, 2, 8, 2, 2, 1};
    // run plans of the recent job sets, most recently used first (a serving
    // loop alternates between a few sets, each with its tile tables and graph)
    std::list<RunPlan> plans;
    RunPlan* cur = nullptr;     // the plan of the current run (front of plans)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int ncu = 256;
    int lds_cu = 160 * 1024;  // LDS bytes per CU
};

static void free_plan(RunPlan& p)
{
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    p.exec = nullptr;
    p.graph = nullptr;
    if (p.djobs) (void)hipFree(p.djobs);
    p.djobs = nullptr;
    for (int4* t : p.dtiles) (void)hipFree(t);
    p.dtiles.clear();
    p.launches.clear();
    p.ids.clear();
}

static constexpr size_t kMaxPlans = 8;

// Frees every cached plan that runs job `id` (all plans when id < 0).
static void drop_plans(tj_batch* b, int id)
{
    for (auto it = b->plans.begin(); it != b->plans.end();) {
        if (id < 0 || std::find(it->ids.begin(), it->ids.end(), id) != it->ids.end()) {
            if (b->cur == &*it) b->cur = nullptr;
            free_plan(*it);
            it = b->plans.erase(it);
        } else {
            ++it;
        }
    }
}

static void free_job(BJob* j)
{
    void* ptrs[] = {j->vol, j->master, j->pairs, j->exps, j->diff, j->err_part, j->prm, j->order, j->err_hist, j->iter};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete j;
}

static size_t job_bytes(const tj_job_desc* d)
{
    const size_t s = storage_bytes(d->storage), a = arith_bytes(d->storage), sp = storage_bytes(proj_storage(*d));
    const size_t nvox = (size_t)d->vol_w * d->vol_h * d->vol_d;
    const size_t npx = (size_t)d->proj_w * d->proj_h * d->n_proj;
    const size_t parts = (size_t)d->n_proj * d->proj_h * fp_col_tiles(d->proj_w);
    const bool half = d->storage == TJ_STORAGE_BF16 || d->storage == TJ_STORAGE_F16;
    const bool master = d->fp32_master && half;
    return nvox * (s + (master ? 4 : 0)) + npx * (sp + a + (half ? 4 : 0)) + parts * 8 +
           (size_t)d->n_proj * (tjb::PRM * 8 + 4) +
           (size_t)std::max(d->max_iterations, 1) * 8 + 256;
}

static int check_desc(const tj_job_desc* d)
{
    if (!d) return bfail(TJ_EINVAL, "null job description");
    if (d->vol_w < 2 || d->vol_h < 1 || d->vol_d < 1 || d->proj_w < 2 || d->proj_h < 1 || d->n_proj < 1)
        return bfail(TJ_EINVAL, "job sizes must be positive (volume width and detector width >= 2)");
    if (d->vol_w > 4096 || d->proj_w > 4096) return bfail(TJ_EUNSUP, "widths above 4096 are not supported");
    if (!storage_bytes(d->storage)) return bfail(TJ_EINVAL, "unknown storage type %d", d->storage);
    if (d->proj_h > d->vol_h)
        return bfail(TJ_EUNSUP, "detector taller than the volume is not supported by the job engine");
    if ((d->vol_h - d->proj_h) % 2 != 0)
        return bfail(TJ_EUNSUP, "volume and detector heights must differ by an even number (integer row shift)");
    if (d->max_iterations < 1) return bfail(TJ_EINVAL, "max_iterations must be >= 1");
    return TJ_OK;
}

// per-projection constants (see the forward/back formulas at the top of the file)
static int job_params(BJob* j)
{
    const tj_job_desc& d = j->d;
    const double pc0 = (d.proj_w - 1) / 2.0, pc1 = (d.proj_h - 1) / 2.0;
    const double vcx = (d.vol_w - 1) / 2.0, vcy = (d.vol_h - 1) / 2.0, vcz = (d.vol_d - 1) / 2.0;
    j->dy = (int)(vcy - pc1);
    j->dyb = (int)(pc1 - vcy);
    j->hprm.assign((size_t)d.n_proj * tjb::PRM, 0.0);
    for (int p = 0; p < d.n_proj; ++p) {
        const double* r = &j->euler[(size_t)9 * p];
        if (r[1] != 0.0 || r[3] != 0.0 || r[5] != 0.0 || r[7] != 0.0 || r[4] != 1.0 || r[0] == 0.0)
            return bfail(TJ_EUNSUP, "projection %d: the job engine needs a tilt about the vertical (y) axis", p);
        const double divx = r[0], zfx = r[2];
        const double dirx = (0.0 - r[2]) / r[0];
        double* q = &j->hprm[(size_t)tjb::PRM * p];
        q[0] = 1.0 / divx;
        q[1] = (-pc0 - zfx * (-1.0 - vcz)) / divx + dirx + vcx;
        q[2] = dirx;
        q[3] = std::sqrt(dirx * dirx + 1.0);
        q[4] = r[0];
        q[5] = r[2];
        q[6] = pc0 - vcx * r[0] - vcz * r[2];
        q[7] = dirx != 0.0 ? 1.0 / dirx : 0.0;
        q[8] = r[2] != 0.0 ? 1.0 / r[2] : 0.0;
    }
    return TJ_OK;
}

// ---------------------------------------------------------------------------
// launch helpers: each launch of a plan resolves, from the storage and the
// tuning, to one kernel instance with its block size and dynamic LDS (Kern);
// the persistent grid is then sized from that kernel's own occupancy.
// ---------------------------------------------------------------------------
// Blocks of `k` one CU holds at once (0 if the runtime cannot say).
static int resident_blocks(const Kern& k)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k.fn, k.block, (size_t)k.lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

// Dynamic LDS of a forward kernel: at most `cap` bytes, and no more than lets
// the CU hold as many blocks as the kernel's registers allow (the forward
// kernels are latency bound: resident blocks count for more than slices per
// slab); never below `need` (one slice of the widest job).
static int fit_lds(Kern k, int cap, int need, int lds_cu)
{
    k.lds = 0;
    const int nb = resident_blocks(k);
    int l = cap;
    if (nb > 0) {
        l = std::min(cap, (lds_cu / nb) & ~1023);
        // allocation granules may round a block's share up: step down to a size
        // at which nb blocks are resident
        for (int m = 0; m < 16 && l - 1024 >= need; ++m) {
            k.lds = l;
            if (resident_blocks(k) >= nb) break;
            l -= 1024;
        }
    }
    return std::max(l, need);
}

template <typename T, typename TE, typename A, int G>
static Kern fp_kern_t(int cap, int max_w, int lds_cu)
{
    using C4 = typename tjb::V4<A>::t;
    Kern k;
    k.fn = (const void*)tjb::k_job_fp<T, TE, A, G>;
    k.lds = fit_lds(k, cap, (max_w + 3) * (int)sizeof(C4), lds_cu);
    k.cells = k.lds / (int)sizeof(C4);
    return k;
}

template <typename T, typename TE, typename A, int G>
static hipError_t allow_lds(int bytes)
{
    return hipFuncSetAttribute((const void*)tjb::k_job_fp<T, TE, A, G>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// Largest dynamic LDS any forward launch may ask for (set once per device,
// before any graph capture).
static const int kMaxFpLds = 152 * 1024;
// Parts (projection groups sharing one slab) of the double-buffered packed
// forward kernel, a 1024-thread block, and rays per thread (two keep it within
// the 128 registers of four waves per SIMD; four spill).
static const int kFpqParts = 4;
static const int kFpqRays = 2;
// The same for the FP32-arithmetic forward kernel k_job_fq (4 rows per tile,
// FQ_CELLS cells per thread and slab).
static const int kFqParts = 4;
static const int kFqRays = 4;

template <typename T, typename TE>
static hipError_t allow_lds2()
{
    hipError_t e = allow_lds<T, TE, float, 2>(kMaxFpLds);
    if (e == hipSuccess) e = allow_lds<T, TE, float, 4>(kMaxFpLds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)tjb::k_job_fq<T, TE, kFqRays, kFqParts>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, kMaxFpLds);
    return e;
}

template <typename T, typename TE>
static hipError_t allow_lds_p()
{
    const void* fns[] = {(const void*)tjb::k_job_fpp<T, TE, 2, 1>, (const void*)tjb::k_job_fpp<T, TE, 4, 1>,
                         (const void*)tjb::k_job_fpp<T, TE, 2, 2>, (const void*)tjb::k_job_fpp<T, TE, 4, 2>,
                         (const void*)tjb::k_job_fpq<T, TE, kFpqRays, kFpqParts>};
    hipError_t e = hipSuccess;
    for (const void* f : fns)
        if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxFpLds);
    return e;
}

static hipError_t allow_fp_lds()
{
    hipError_t e = allow_lds2<float, float>();
    if (e == hipSuccess) e = allow_lds2<Bf16, Bf16>();
    if (e == hipSuccess) e = allow_lds2<Bf16, float>();
    if (e == hipSuccess) e = allow_lds2<_Float16, _Float16>();
    if (e == hipSuccess) e = allow_lds2<_Float16, float>();
    if (e == hipSuccess) e = allow_lds<double, double, double, 2>(kMaxFpLds);
    if (e == hipSuccess) e = allow_lds_p<Bf16, Bf16>();
    if (e == hipSuccess) e = allow_lds_p<Bf16, float>();
    if (e == hipSuccess) e = allow_lds_p<_Float16, _Float16>();
    if (e == hipSuccess) e = allow_lds_p<_Float16, float>();
    return e;
}

static int fp_lds_cap(const tj_batch_tuning& tu) { return std::min(std::max(tu.fp_lds_kb, 16), 152) * 1024; }

// k_job_fq takes launches whose widest slab row (W + 1 cells) fits its
// threads' FQ_CELLS cells each
static bool fq_ok(int max_w, const tj_batch_tuning& tu)
{
    return tu.fp_shared && max_w + 1 <= tjb::FQ_CELLS * kFqParts * tjb::NT;
}

template <typename T, typename TE>
static Kern fp_kern_s(const tj_batch_tuning& tu, int max_w, int lds_cu)
{
    if (fq_ok(max_w, tu)) {
        Kern k;
        k.block = kFqParts * tjb::NT;
        k.fn = (const void*)tjb::k_job_fq<T, TE, kFqRays, kFqParts>;
        // one 1024-thread block fills a CU: all the LDS a block may have
        k.lds = fit_lds(k, kMaxFpLds, 2 * (max_w + 3) * (int)sizeof(float4), lds_cu);
        k.cells = k.lds / (int)sizeof(float4);
        return k;
    }
    if (tu.fp_rays <= 2) return fp_kern_t<T, TE, float, 2>(fp_lds_cap(tu), max_w, lds_cu);
    return fp_kern_t<T, TE, float, 4>(fp_lds_cap(tu), max_w, lds_cu);
}

static bool use_packed(const tj_batch_tuning& tu, int storage)
{
    return tu.fp_packed && (storage == TJ_STORAGE_BF16 || storage == TJ_STORAGE_F16);
}

// 256-column tiles per packed forward block: 1 or 2
static int fp_cols_of(const tj_batch_tuning& tu, int storage)
{
    return (use_packed(tu, storage) && tu.fp_cols >= 2) ? 2 : 1;
}

// projections per forward tile: 2 or 4
static int fp_rays_of(const tj_batch_tuning& tu, int storage)
{
    if (storage == TJ_STORAGE_F64) return 2;
    return tu.fp_rays <= 2 ? 2 : 4;
}

// packed-pair forward kernel for 16-bit storage (8 rows per tile); one slice
// of the widest job needs 2 (W + 2) cells plus the padding of the odd half
// Cells of one slice of the packed slab (even cells, the odd half from hb on).
static int fpp_slice_cells(int w)
{
    const int wc = w + 2, hb = (((wc + 1) / 2 + 3) & ~7) + 4;
    return hb + wc / 2;
}

// The double-buffered packed forward kernel takes a launch whose jobs all have
// an even width of at most 2 * (its threads) columns (one column pair per
// thread and slab) and whose widest slice fits the LDS twice.
static bool fpq_ok(const std::vector<BJob*>& js, const tj_batch_tuning& tu, int storage)
{
    if (!use_packed(tu, storage) || tu.fp_packed != 2) return false;
    for (const BJob* j : js) {
        const int w = j->d.vol_w;
        if ((w & 1) || w / 2 > kFpqParts * tjb::NT || fpp_slice_cells(w) > tjb::FPQ_PLANE) return false;
    }
    return true;
}

template <typename T, typename TE>
static Kern fpp_kern_s(const tj_batch_tuning& tu, int max_w, int lds_cu, bool quad)
{
    if (quad) {
        Kern k;
        k.block = kFpqParts * tjb::NT;
        k.fn = (const void*)tjb::k_job_fpq<T, TE, kFpqRays, kFpqParts>;
        // one 1024-thread block fills a CU: two buffers of two fixed planes
        // (zc = 4 slices at W = 512, 2 at W = 1024)
        static_assert(4 * tjb::FPQ_PLANE * (int)sizeof(uint4) <= kMaxFpLds, "k_job_fpq slab buffers");
        k.lds = 4 * tjb::FPQ_PLANE * (int)sizeof(uint4);
        k.cells = k.lds / (int)sizeof(uint4);
        return k;
    }
    const bool two = fp_cols_of(tu, TJ_STORAGE_BF16) == 2, g2 = fp_rays_of(tu, TJ_STORAGE_BF16) == 2;
    Kern k;
    k.block = two ? 2 * tjb::NT : tjb::NT;
    if (two) k.fn = g2 ? (const void*)tjb::k_job_fpp<T, TE, 2, 2> : (const void*)tjb::k_job_fpp<T, TE, 4, 2>;
    else k.fn = g2 ? (const void*)tjb::k_job_fpp<T, TE, 2, 1> : (const void*)tjb::k_job_fpp<T, TE, 4, 1>;
    k.lds = fit_lds(k, fp_lds_cap(tu), 2 * (max_w + 16) * (int)sizeof(uint4), lds_cu);
    k.cells = k.lds / (int)sizeof(uint4);
    return k;
}

static int fp_rows_of(const tj_batch_tuning& tu, int storage) { return use_packed(tu, storage) ? 8 : 4; }

static Kern fp_kern(int storage, int pstorage, const tj_batch_tuning& tu, int max_w, int lds_cu, bool quad)
{
    const bool pf = pstorage == TJ_STORAGE_F32;
    if (use_packed(tu, storage)) {
        if (storage == TJ_STORAGE_BF16)
            return pf ? fpp_kern_s<Bf16, float>(tu, max_w, lds_cu, quad)
                      : fpp_kern_s<Bf16, Bf16>(tu, max_w, lds_cu, quad);
        return pf ? fpp_kern_s<_Float16, float>(tu, max_w, lds_cu, quad)
                  : fpp_kern_s<_Float16, _Float16>(tu, max_w, lds_cu, quad);
    }
    switch (storage) {
    case TJ_STORAGE_F32: return fp_kern_s<float, float>(tu, max_w, lds_cu);
    case TJ_STORAGE_BF16:
        return pf ? fp_kern_s<Bf16, float>(tu, max_w, lds_cu) : fp_kern_s<Bf16, Bf16>(tu, max_w, lds_cu);
    case TJ_STORAGE_F16:
        return pf ? fp_kern_s<_Float16, float>(tu, max_w, lds_cu) : fp_kern_s<_Float16, _Float16>(tu, max_w, lds_cu);
    default: return fp_kern_t<double, double, double, 2>(128 * 1024, max_w, lds_cu);
    }
}

// backprojection shapes: (rows, slices, images per LDS pass)
struct BpShape {
    int r, kc, g;
};
// (rows, slices) from the tuning: rows 2, 4 or 8; slices 4, 8 or 16
static BpShape bp_shape_of(const tj_batch_tuning& tu, int storage)
{
    if (storage == TJ_STORAGE_F64) return {2, 8, 2};
    if (tu.bp_rows <= 2) return tu.bp_zchunk >= 16 ? BpShape{2, 16, 2} : BpShape{2, 8, 4};
    if (tu.bp_rows >= 8) return tu.bp_zchunk >= 16 ? BpShape{8, 16, 2} : BpShape{8, 8, 2};
    if (tu.bp_zchunk <= 4) return {4, 4, 4};
    return tu.bp_zchunk >= 16 ? BpShape{4, 16, 4} : BpShape{4, 8, 4};
}

template <typename T, int MODE>
static const void* bp_fn_s(const BpShape& s)
{
    if (s.r == 2 && s.kc == 16) return (const void*)tjb::k_job_bp<T, float, 2, 16, 2, MODE>;
    if (s.r == 2) return (const void*)tjb::k_job_bp<T, float, 2, 8, 4, MODE>;
    if (s.r == 8 && s.kc == 16) return (const void*)tjb::k_job_bp<T, float, 8, 16, 2, MODE>;
    if (s.r == 8) return (const void*)tjb::k_job_bp<T, float, 8, 8, 2, MODE>;
    if (s.kc == 4) return (const void*)tjb::k_job_bp<T, float, 4, 4, 4, MODE>;
    if (s.kc == 16) return (const void*)tjb::k_job_bp<T, float, 4, 16, 4, MODE>;
    return (const void*)tjb::k_job_bp<T, float, 4, 8, 4, MODE>;
}

// 16-bit storage always takes the FP32-cell-position form (mode bit 2)
template <typename T>
static const void* bp_fn_m(int mode, const BpShape& s)
{
    switch (mode) {
    case 0: return bp_fn_s<T, 0>(s);
    case 1: return bp_fn_s<T, 1>(s);
    case 4: return bp_fn_s<T, 4>(s);
    case 5: return bp_fn_s<T, 5>(s);
    case 6: return bp_fn_s<T, 6>(s);
    default: return bp_fn_s<T, 7>(s);
    }
}

// mode: bit 0 positivity, bit 1 FP32 master volume (16-bit storage only)
static Kern bp_kern(int storage, int mode, const tj_batch_tuning& tu)
{
    const BpShape s = bp_shape_of(tu, storage);
    Kern k;
    switch (storage) {
    case TJ_STORAGE_F32: k.fn = bp_fn_m<float>(mode & 1, s); break;
    case TJ_STORAGE_BF16: k.fn = bp_fn_m<Bf16>(mode | 4, s); break;
    case TJ_STORAGE_F16: k.fn = bp_fn_m<_Float16>(mode | 4, s); break;
    default:
        k.fn = (mode & 1) ? (const void*)tjb::k_job_bp<double, double, 2, 8, 2, 1>
                          : (const void*)tjb::k_job_bp<double, double, 2, 8, 2, 0>;
    }
    return k;
}

// One persistent launch of a resolved kernel over a tile table.
static hipError_t launch_kern(hipStream_t st, const Kern& k, int grid, const TjJob* jobs, const int4* tiles,
                              int ntiles)
{
    int cells = k.cells;
    void* args[] = {(void*)&jobs, (void*)&tiles, (void*)&ntiles, (void*)&cells};
    (void)hipGetLastError();
    return hipLaunchKernel(k.fn, dim3(grid), dim3(k.block), args, (size_t)k.lds, st);
}

// ---------------------------------------------------------------------------
// tile tables.  Tiles that read the same data (the same volume rows for the
// forward projection, the same image rows for the backprojection) form a
// group; groups are dealt, largest first, to the 8 XCD queues, and the queues
// are interleaved so that tile t runs on XCD t % 8 (blocks are dealt to XCDs
// round-robin and each block takes tiles t = b, b + grid, ... with grid a
// multiple of 8).  Short queues are padded with empty tiles.
// ---------------------------------------------------------------------------
struct Group {
    double cost;
    std::vector<int4> tiles;
};

static std::vector<int4> interleave(std::vector<Group>& groups)
{
    std::stable_sort(groups.begin(), groups.end(), [](const Group& a, const Group& b) { return a.cost > b.cost; });
    std::vector<int4> q[8];
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (Group& g : groups) {
        int best = 0;
        for (int x = 1; x < 8; ++x)
            if (load[x] < load[best]) best = x;
        load[best] += g.cost;
        q[best].insert(q[best].end(), g.tiles.begin(), g.tiles.end());
    }
    size_t len = 0;
    for (auto& v : q) len = std::max(len, v.size());
    std::vector<int4> out(len * 8, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < 8; ++x)
        for (size_t m = 0; m < q[x].size(); ++m) out[m * 8 + x] = q[x][m];
    return out;
}

// CT: 256-column tiles per block (the tile's column index counts blocks)
static std::vector<int4> fp_tiles(const std::vector<BJob*>& js, const std::vector<int>& idx, int G, int RB,
                                  int CT)
{
    std::vector<Group> groups;
    for (size_t m = 0; m < js.size(); ++m) {
        const tj_job_desc& d = js[m]->d;
        const int nct = (fp_col_tiles(d.proj_w) + CT - 1) / CT, nrb = (d.proj_h + RB - 1) / RB;
        for (int rb = 0; rb < nrb; ++rb) {
            Group g;
            g.cost = (double)d.vol_d * ((double)d.vol_w + (double)tjb::NT * CT * G) * ((d.n_proj + G - 1) / G) * nct;
            for (int p0 = 0; p0 < d.n_proj; p0 += G)
                for (int ct = 0; ct < nct; ++ct) g.tiles.push_back(make_int4(idx[m], p0, rb, ct));
            groups.push_back(std::move(g));
        }
    }
    return interleave(groups);
}

static std::vector<int4> bp_tiles(const std::vector<BJob*>& js, const std::vector<int>& idx, const BpShape& s)
{
    std::vector<Group> groups;
    for (size_t m = 0; m < js.size(); ++m) {
        const tj_job_desc& d = js[m]->d;
        const int nxt = (d.vol_w + tjb::NT - 1) / tjb::NT, nrb = (d.vol_h + s.r - 1) / s.r;
        const int nkb = (d.vol_d + s.kc - 1) / s.kc;
        for (int rb = 0; rb < nrb; ++rb) {
            Group g;
            g.cost = (double)d.n_proj * nxt * nkb * (s.kc * s.r * (double)tjb::NT + (tjb::NT + s.kc) * s.r);
            for (int xt = 0; xt < nxt; ++xt)
                for (int kb = 0; kb < nkb; ++kb) g.tiles.push_back(make_int4(idx[m], xt, rb, kb));
            groups.push_back(std::move(g));
        }
    }
    return interleave(groups);
}

// Persistent grid of a launch: blocks_per_cu blocks per CU, at most as many as
// the CU holds at once (blocks beyond those would run as a second, partial
// round and leave the first round's CUs idle at the end), at most one per tile,
// and a multiple of 8 (tile t runs on XCD t % 8).
static int grid_for(const tj_batch* b, int ntiles, const Kern& k)
{
    int per_cu = std::max(1, std::min(b->tune.blocks_per_cu, 16));
    const int nb = resident_blocks(k);
    if (nb > 0) per_cu = std::min(per_cu, nb);
    int g = std::min(ntiles, per_cu * b->ncu);
    g = ((g + 7) / 8) * 8;
    return std::max(g, 8);
}

static TjJob dev_desc(const BJob* j)
{
    TjJob t{};
    t.vol = j->vol;
    t.master = j->master;
    t.pairs = j->pairs;
    t.exps = j->exps;
    t.diff = j->diff;
    t.err_part = j->err_part;
    t.prm = j->prm;
    t.order = j->order;
    t.err_hist = j->err_hist;
    t.iter = j->iter;
    t.W = j->d.vol_w;
    t.H = j->d.vol_h;
    t.D = j->d.vol_d;
    t.pw = j->d.proj_w;
    t.ph = j->d.proj_h;
    t.n = j->d.n_proj;
    t.dy = j->dy;
    t.dyb = j->dyb;
    t.nct = fp_col_tiles(j->d.proj_w);
    t.hist_cap = std::max(j->d.max_iterations, 1);
    // TomoJ casts lambda/update to float (TomoReconstruction2.java:646); the
    // FP64 twin keeps it in double
    t.factor = j->d.storage == TJ_STORAGE_F64 ? j->d.relaxation / j->d.n_proj
                                              : (double)(float)(j->d.relaxation / j->d.n_proj);
    return t;
}

// Frees the cached plan of exactly this job set (stale tuning).
static void drop_plans_of_set(tj_batch* b, const std::vector<int>& ids)
{
    for (auto it = b->plans.begin(); it != b->plans.end();) {
        if (it->ids == ids) {
            if (b->cur == &*it) b->cur = nullptr;
            free_plan(*it);
            it = b->plans.erase(it);
        } else {
            ++it;
        }
    }
}

// Builds the launches of one iteration over `ids`: per storage type a forward
// launch, one error launch over every job, per (storage, positivity) a back
// launch.  Jobs are ordered by storage in the descriptor table.
static int build_plan(tj_batch* b, const std::vector<int>& ids)
{
    drop_plans_of_set(b, ids);
    while (b->plans.size() >= kMaxPlans) {
        if (b->cur == &b->plans.back()) b->cur = nullptr;
        free_plan(b->plans.back());
        b->plans.pop_back();
    }
    b->plans.emplace_front();
    b->cur = &b->plans.front();
    RunPlan& p = *b->cur;
    std::vector<BJob*> js;
    for (int id : ids) {
        auto it = b->jobs.find(id);
        if (it == b->jobs.end()) return bfail(TJ_EINVAL, "unknown job %d", id);
        js.push_back(it->second);
    }
    std::stable_sort(js.begin(), js.end(), [](const BJob* a, const BJob* c) { return a->d.storage < c->d.storage; });
    std::vector<TjJob> hj;
    for (BJob* j : js) hj.push_back(dev_desc(j));
    TB_HIP(hipMalloc((void**)&p.djobs, sizeof(TjJob) * hj.size()));
    TB_HIP(hipMemcpy(p.djobs, hj.data(), sizeof(TjJob) * hj.size(), hipMemcpyHostToDevice));
    auto upload = [&](const std::vector<int4>& v, int4** out) -> int {
        *out = nullptr;
        if (v.empty()) return TJ_OK;
        TB_HIP(hipMalloc((void**)out, sizeof(int4) * v.size()));
        p.dtiles.push_back(*out);
        TB_HIP(hipMemcpy(*out, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice));
        return TJ_OK;
    };
    std::vector<RunPlan::L> fwd, back;
    for (int s = 0; s <= TJ_STORAGE_F64; ++s) {
        std::vector<BJob*> sj;
        std::vector<int> sidx;
        for (size_t m = 0; m < js.size(); ++m)
            if (js[m]->d.storage == s) {
                sj.push_back(js[m]);
                sidx.push_back((int)m);
            }
        if (sj.empty()) continue;
        for (int ps = 0; ps <= TJ_STORAGE_F64; ++ps) {
            std::vector<BJob*> fj;
            std::vector<int> fidx;
            for (size_t m = 0; m < sj.size(); ++m)
                if (proj_storage(sj[m]->d) == ps) {
                    fj.push_back(sj[m]);
                    fidx.push_back(sidx[m]);
                }
            if (fj.empty()) continue;
            // the double-buffered kernels' tiles step over all their parts'
            // projections and cover one 256-column tile each
            int max_w = 0;
            for (BJob* j : fj) max_w = std::max(max_w, j->d.vol_w);
            const bool quad = fpq_ok(fj, b->tune, s);
            const bool shared = !use_packed(b->tune, s) && s != TJ_STORAGE_F64 && fq_ok(max_w, b->tune);
            const int fg = quad ? kFpqRays * kFpqParts : shared ? kFqRays * kFqParts : fp_rays_of(b->tune, s);
            std::vector<int4> ft =
                fp_tiles(fj, fidx, fg, fp_rows_of(b->tune, s), (quad || shared) ? 1 : fp_cols_of(b->tune, s));
            int4* dft = nullptr;
            if (int rc = upload(ft, &dft)) return rc;
            const Kern k = fp_kern(s, ps, b->tune, max_w, b->lds_cu, quad);
            fwd.push_back({0, s, max_w, (int)ft.size(), grid_for(b, (int)ft.size(), k), dft, 0, 0, ps, k});
        }
        for (int mode = 0; mode <= 3; ++mode) {
            std::vector<BJob*> pj;
            std::vector<int> pidx;
            for (size_t m = 0; m < sj.size(); ++m)
                if (((sj[m]->d.positivity != 0 ? 1 : 0) | (sj[m]->master ? 2 : 0)) == mode) {
                    pj.push_back(sj[m]);
                    pidx.push_back(sidx[m]);
                }
            if (pj.empty()) continue;
            std::vector<int4> bt = bp_tiles(pj, pidx, bp_shape_of(b->tune, s));
            int4* dbt = nullptr;
            if (int rc = upload(bt, &dbt)) return rc;
            const Kern k = bp_kern(s, mode, b->tune);
            back.push_back({2, s, mode, (int)bt.size(), grid_for(b, (int)bt.size(), k), dbt, 0, 0, s, k});
        }
    }
    p.launches = fwd;
    p.launches.push_back({1, 0, 0, 0, (int)js.size(), nullptr, 0, (int)js.size(), 0});
    {
        // 16-bit jobs are contiguous in the descriptor table (sorted by storage)
        int h0 = -1, hn = 0;
        for (size_t m = 0; m < js.size(); ++m)
            if (js[m]->pairs) {
                if (h0 < 0) h0 = (int)m;
                ++hn;
            }
        if (hn > 0) p.launches.push_back({3, 0, 0, 0, 0, nullptr, h0, hn, 0});
    }
    p.launches.insert(p.launches.end(), back.begin(), back.end());
    p.ids = ids;
    p.tune = b->tune;
    return TJ_OK;
}

static hipError_t enqueue_launch(tj_batch* b, const RunPlan::L& l)
{
    const RunPlan& p = *b->cur;
    if (l.kind == 0 || l.kind == 2) return launch_kern(b->stream, l.k, l.grid, p.djobs, l.tiles, l.ntiles);
    (void)hipGetLastError();
    if (l.kind == 3) {
        tjb::k_job_pairs<<<dim3(2048, l.njobs), tjb::NT, 0, b->stream>>>(p.djobs + l.job0);
        return hipGetLastError();
    }
    tjb::k_job_err<<<l.njobs, tjb::NTE, 0, b->stream>>>(p.djobs + l.job0, l.njobs);
    return hipGetLastError();
}

static hipError_t enqueue_iterations(tj_batch* b, int iterations)
{
    for (int it = 0; it < iterations; ++it)
        for (const RunPlan::L& l : b->cur->launches) {
            hipError_t e = enqueue_launch(b, l);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

// Makes the cached plan of this job set (built with the current tuning) the
// current one; false when there is none.
static bool plan_matches(tj_batch* b, const std::vector<int>& ids)
{
    for (auto it = b->plans.begin(); it != b->plans.end(); ++it) {
        if (it->ids == ids && memcmp(&it->tune, &b->tune, sizeof(tj_batch_tuning)) == 0) {
            b->plans.splice(b->plans.begin(), b->plans, it);
            b->cur = &b->plans.front();
            return true;
        }
    }
    return false;
}

template <typename TO>
static hipError_t convert_to(hipStream_t st, int from, const void* src, TO* dst, size_t n)
{
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    switch (from) {
    case TJ_STORAGE_F32: tjb::k_convert<float, TO><<<blocks, 256, 0, st>>>((const float*)src, dst, n); break;
    case TJ_STORAGE_BF16: tjb::k_convert<Bf16, TO><<<blocks, 256, 0, st>>>((const Bf16*)src, dst, n); break;
    case TJ_STORAGE_F16: tjb::k_convert<_Float16, TO><<<blocks, 256, 0, st>>>((const _Float16*)src, dst, n); break;
    default: tjb::k_convert<double, TO><<<blocks, 256, 0, st>>>((const double*)src, dst, n); break;
    }
    return hipGetLastError();
}

static hipError_t convert_any(hipStream_t st, int from, const void* src, int to, void* dst, size_t n)
{
    switch (to) {
    case TJ_STORAGE_F32: return convert_to(st, from, src, (float*)dst, n);
    case TJ_STORAGE_BF16: return convert_to(st, from, src, (Bf16*)dst, n);
    case TJ_STORAGE_F16: return convert_to(st, from, src, (_Float16*)dst, n);
    default: return convert_to(st, from, src, (double*)dst, n);
    }
}

static BJob* find_job(tj_batch* b, int id)
{
    auto it = b->jobs.find(id);
    return it == b->jobs.end() ? nullptr : it->second;
}

template <typename T>
static hipError_t put_volume(hipStream_t st, const float* src, void* vol, int W, int H, int D)
{
    tjb::k_zyx_to_job<T><<<dim3((W + 255) / 256, H, D), 256, 0, st>>>(src, (T*)vol, W, H, D);
    return hipGetLastError();
}
template <typename T>
static hipError_t take_volume(hipStream_t st, const void* vol, float* dst, int W, int H, int D)
{
    tjb::k_job_to_zyx<T><<<dim3((W + 255) / 256, H, D), 256, 0, st>>>((const T*)vol, dst, W, H, D);
    return hipGetLastError();
}

extern "C" {

size_t tj_job_bytes(const tj_job_desc* d)
{
    if (check_desc(d)) return 0;
    return job_bytes(d);
}

int tj_batch_open(int device, size_t hbm_budget, tj_batch** out)
{
    if (!out) return bfail(TJ_EINVAL, "tj_batch_open: null output");
    *out = nullptr;
    int ndev = 0;
    TB_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return bfail(TJ_EINVAL, "tj_batch_open: no HIP device %d", device);
    TB_HIP(hipSetDevice(device));
    tj_batch* b = new tj_batch();
    b->device = device;
    b->budget = hbm_budget;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->ncu = prop.multiProcessorCount;
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess &&
        lds_cu > 0)
        b->lds_cu = lds_cu;
    if (allow_fp_lds() != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        delete b;
        return bfail(TJ_EDEVICE, "tj_batch_open: cannot create stream/events");
    }
    *out = b;
    return TJ_OK;
}

int tj_batch_close(tj_batch* b)
{
    if (!b) return TJ_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_plans(b, -1);
    for (auto& kv : b->jobs) free_job(kv.second);
    (void)hipEventDestroy(b->ev0);
    (void)hipEventDestroy(b->ev1);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return TJ_OK;
}

size_t tj_batch_used(const tj_batch* b) { return b ? b->used : 0; }
size_t tj_batch_budget(const tj_batch* b) { return b ? b->budget : 0; }

int tj_batch_add_job(tj_batch* b, const tj_job_desc* d, const double* euler, int* job)
{
    if (!b || !euler || !job) return bfail(TJ_EINVAL, "tj_batch_add_job: null argument");
    if (int rc = check_desc(d)) return rc;
    const size_t need = job_bytes(d);
    if (b->used + need > b->budget)
        return bfail(TJ_ENOMEM, "tj_batch_add_job: %zu bytes needed, %zu of %zu in use", need, b->used, b->budget);
    TB_HIP(hipSetDevice(b->device));
    BJob* j = new BJob();
    j->d = *d;
    j->euler.assign(euler, euler + (size_t)9 * d->n_proj);
    if (int rc = job_params(j)) {
        delete j;
        return rc;
    }
    const size_t s = storage_bytes(d->storage), a = arith_bytes(d->storage), sp = storage_bytes(proj_storage(*d));
    const size_t nvox = (size_t)d->vol_w * d->vol_h * d->vol_d;
    const size_t npx = (size_t)d->proj_w * d->proj_h * d->n_proj;
    const size_t parts = (size_t)d->n_proj * d->proj_h * fp_col_tiles(d->proj_w);
    hipError_t e = hipMalloc(&j->vol, nvox * s);
    if (e == hipSuccess && d->fp32_master && (d->storage == TJ_STORAGE_BF16 || d->storage == TJ_STORAGE_F16)) {
        e = hipMalloc((void**)&j->master, nvox * 4);
        if (e == hipSuccess) e = hipMemsetAsync(j->master, 0, nvox * 4, b->stream);
    }
    if (e == hipSuccess) e = hipMalloc((void**)&j->exps, npx * sp);
    if (e == hipSuccess) e = hipMalloc(&j->diff, npx * a);
    if (e == hipSuccess && (d->storage == TJ_STORAGE_BF16 || d->storage == TJ_STORAGE_F16))
        e = hipMalloc((void**)&j->pairs, npx * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&j->err_part, parts * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&j->prm, j->hprm.size() * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&j->err_hist, (size_t)d->max_iterations * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&j->iter, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(j->prm, j->hprm.data(), j->hprm.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // SIRT adds the images in orderShuffleART order from image 0 (TomoJ draws
        // the first image at random; the plain path takes first_index likewise)
        std::vector<int> ord(d->n_proj);
        tj_order_shuffle(j->euler.data(), d->n_proj, 0, ord.data());
        e = hipMalloc((void**)&j->order, sizeof(int) * d->n_proj);
        if (e == hipSuccess) e = hipMemcpy(j->order, ord.data(), sizeof(int) * d->n_proj, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemsetAsync(j->vol, 0, nvox * s, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(j->exps, 0, npx * sp, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(j->iter, 0, sizeof(int), b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        free_job(j);
        return bfail(TJ_ENOMEM, "tj_batch_add_job: device allocation failed: %s", hipGetErrorString(e));
    }
    j->id = b->next_id++;
    j->bytes = need;
    b->used += need;
    b->jobs[j->id] = j;
    *job = j->id;
    return TJ_OK;
}

int tj_batch_remove_job(tj_batch* b, int id)
{
    if (!b) return bfail(TJ_EINVAL, "null batch");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "tj_batch_remove_job: unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    TB_HIP(hipStreamSynchronize(b->stream));
    drop_plans(b, id);
    b->used -= j->bytes;
    b->jobs.erase(id);
    free_job(j);
    return TJ_OK;
}

int tj_batch_set_projections(tj_batch* b, int id, const float* stack)
{
    if (!b || !stack) return bfail(TJ_EINVAL, "tj_batch_set_projections: null argument");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    const size_t npx = (size_t)j->d.proj_w * j->d.proj_h * j->d.n_proj;
    float* tmp = nullptr;
    TB_HIP(hipMalloc((void**)&tmp, npx * 4));
    hipError_t e = hipMemcpyAsync(tmp, stack, npx * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = convert_any(b->stream, TJ_STORAGE_F32, tmp, proj_storage(j->d), j->exps, npx);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "tj_batch_set_projections: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_batch_synth_job(tj_batch* b, int id, const double* ell, int k, double noise_rel, uint64_t seed)
{
    if (!b || (!ell && k > 0) || k < 0) return bfail(TJ_EINVAL, "tj_batch_synth_job: bad argument");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    const tj_job_desc& d = j->d;
    const size_t npx = (size_t)d.proj_w * d.proj_h * d.n_proj;
    std::vector<TjProj> hp(d.n_proj);
    for (int p = 0; p < d.n_proj; ++p) {
        memset(&hp[p], 0, sizeof(TjProj));
        for (int q = 0; q < 9; ++q) hp[p].e[q] = j->euler[(size_t)9 * p + q];
    }
    TjProj* dp = nullptr;
    double *dell = nullptr, *clean = nullptr, *part = nullptr;
    float* tmp = nullptr;
    const int npart = 256;
    hipError_t e = hipMalloc((void**)&dp, sizeof(TjProj) * d.n_proj);
    if (e == hipSuccess) e = hipMalloc((void**)&dell, sizeof(double) * 16 * std::max(k, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&clean, sizeof(double) * npx);
    if (e == hipSuccess) e = hipMalloc((void**)&part, sizeof(double) * npart);
    if (e == hipSuccess) e = hipMalloc((void**)&tmp, sizeof(float) * npx);
    if (e == hipSuccess) e = hipMemcpy(dp, hp.data(), sizeof(TjProj) * d.n_proj, hipMemcpyHostToDevice);
    if (e == hipSuccess && k > 0) e = hipMemcpy(dell, ell, sizeof(double) * 16 * k, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = tjk::phantom_project(b->stream, dell, k, dp, d.n_proj, d.proj_w, d.proj_h, (d.proj_w - 1) / 2.0,
                                 (d.proj_h - 1) / 2.0, clean);
    if (e == hipSuccess) e = tjk::add_noise(b->stream, clean, npx, seed, noise_rel, part, npart, tmp);
    if (e == hipSuccess) e = convert_any(b->stream, TJ_STORAGE_F32, tmp, proj_storage(d), j->exps, npx);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    void* ptrs[] = {dp, dell, clean, part, tmp};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "tj_batch_synth_job: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_batch_get_projection(tj_batch* b, int id, int index, float* out)
{
    if (!b || !out) return bfail(TJ_EINVAL, "null argument");
    BJob* j = find_job(b, id);
    if (!j || index < 0 || index >= j->d.n_proj) return bfail(TJ_EINVAL, "bad job or projection index");
    TB_HIP(hipSetDevice(b->device));
    const size_t npx = (size_t)j->d.proj_w * j->d.proj_h;
    float* tmp = nullptr;
    TB_HIP(hipMalloc((void**)&tmp, npx * 4));
    const int ps = proj_storage(j->d);
    const char* src = (const char*)j->exps + npx * index * storage_bytes(ps);
    hipError_t e = convert_any(b->stream, ps, src, TJ_STORAGE_F32, tmp, npx);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, npx * 4, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "tj_batch_get_projection: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_batch_get_difference(tj_batch* b, int id, int index, double* out)
{
    if (!b || !out) return bfail(TJ_EINVAL, "null argument");
    BJob* j = find_job(b, id);
    if (!j || index < 0 || index >= j->d.n_proj) return bfail(TJ_EINVAL, "bad job or projection index");
    TB_HIP(hipSetDevice(b->device));
    const size_t npx = (size_t)j->d.proj_w * j->d.proj_h;
    double* tmp = nullptr;
    TB_HIP(hipMalloc((void**)&tmp, npx * 8));
    const int as = j->d.storage == TJ_STORAGE_F64 ? TJ_STORAGE_F64 : TJ_STORAGE_F32;
    const char* src = (const char*)j->diff + npx * index * arith_bytes(j->d.storage);
    hipError_t e = convert_any(b->stream, as, src, TJ_STORAGE_F64, tmp, npx);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, npx * 8, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "tj_batch_get_difference: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_batch_set_volume(tj_batch* b, int id, const float* zyx)
{
    if (!b || !zyx) return bfail(TJ_EINVAL, "null argument");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    const int W = j->d.vol_w, H = j->d.vol_h, D = j->d.vol_d;
    const size_t nvox = (size_t)W * H * D;
    float* tmp = nullptr;
    TB_HIP(hipMalloc((void**)&tmp, nvox * 4));
    hipError_t e = hipMemcpyAsync(tmp, zyx, nvox * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
        switch (j->d.storage) {
        case TJ_STORAGE_F32: e = put_volume<float>(b->stream, tmp, j->vol, W, H, D); break;
        case TJ_STORAGE_BF16: e = put_volume<Bf16>(b->stream, tmp, j->vol, W, H, D); break;
        case TJ_STORAGE_F16: e = put_volume<_Float16>(b->stream, tmp, j->vol, W, H, D); break;
        default: e = put_volume<double>(b->stream, tmp, j->vol, W, H, D); break;
        }
    }
    if (e == hipSuccess && j->master) e = put_volume<float>(b->stream, tmp, j->master, W, H, D);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "tj_batch_set_volume: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_batch_get_volume(tj_batch* b, int id, float* zyx)
{
    if (!b || !zyx) return bfail(TJ_EINVAL, "null argument");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    const int W = j->d.vol_w, H = j->d.vol_h, D = j->d.vol_d;
    const size_t nvox = (size_t)W * H * D;
    float* tmp = nullptr;
    TB_HIP(hipMalloc((void**)&tmp, nvox * 4));
    hipError_t e;
    switch (j->master ? TJ_STORAGE_F32 : j->d.storage) {
    case TJ_STORAGE_F32: e = take_volume<float>(b->stream, j->master ? (const void*)j->master : j->vol, tmp, W, H, D); break;
    case TJ_STORAGE_BF16: e = take_volume<Bf16>(b->stream, j->vol, tmp, W, H, D); break;
    case TJ_STORAGE_F16: e = take_volume<_Float16>(b->stream, j->vol, tmp, W, H, D); break;
    default: e = take_volume<double>(b->stream, j->vol, tmp, W, H, D); break;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(zyx, tmp, nvox * 4, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "tj_batch_get_volume: %s", hipGetErrorString(e));
    return TJ_OK;
}

int tj_batch_get_volume_f64(tj_batch* b, int id, double* zyx)
{
    if (!b || !zyx) return bfail(TJ_EINVAL, "null argument");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    if (j->d.storage != TJ_STORAGE_F64) return bfail(TJ_EINVAL, "job %d does not store FP64", id);
    TB_HIP(hipSetDevice(b->device));
    const int W = j->d.vol_w, H = j->d.vol_h, D = j->d.vol_d;
    // [y][z][x] -> host, reordered on the host
    std::vector<double> tmp((size_t)W * H * D);
    TB_HIP(hipMemcpy(tmp.data(), j->vol, tmp.size() * 8, hipMemcpyDeviceToHost));
    for (int y = 0; y < H; ++y)
        for (int z = 0; z < D; ++z)
            memcpy(zyx + ((size_t)z * H + y) * W, tmp.data() + ((size_t)y * D + z) * W, (size_t)W * 8);
    return TJ_OK;
}

int tj_batch_convert(tj_batch* b, int id, int storage)
{
    if (!b) return bfail(TJ_EINVAL, "null batch");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    if (!storage_bytes(storage)) return bfail(TJ_EINVAL, "unknown storage type %d", storage);
    if (storage == j->d.storage) return TJ_OK;
    tj_job_desc nd = j->d;
    nd.storage = storage;
    const size_t need = job_bytes(&nd);
    if (b->used - j->bytes + need > b->budget)
        return bfail(TJ_ENOMEM, "tj_batch_convert: %zu bytes needed over the budget", need);
    TB_HIP(hipSetDevice(b->device));
    TB_HIP(hipStreamSynchronize(b->stream));
    const size_t nvox = (size_t)nd.vol_w * nd.vol_h * nd.vol_d;
    const size_t npx = (size_t)nd.proj_w * nd.proj_h * nd.n_proj;
    void *vol = nullptr, *exps = nullptr, *diff = nullptr;
    uint32_t* pairs = nullptr;
    const bool half_new = storage == TJ_STORAGE_BF16 || storage == TJ_STORAGE_F16;
    hipError_t e = hipMalloc(&vol, nvox * storage_bytes(storage));
    if (e == hipSuccess && half_new && !j->pairs) e = hipMalloc((void**)&pairs, npx * 4);
    const int ps_old = proj_storage(j->d), ps_new = proj_storage(nd);
    if (e == hipSuccess) e = hipMalloc(&exps, npx * storage_bytes(ps_new));
    if (e == hipSuccess && arith_bytes(storage) != arith_bytes(j->d.storage))
        e = hipMalloc(&diff, npx * arith_bytes(storage));
    if (e == hipSuccess) {
        if (j->master) e = convert_any(b->stream, TJ_STORAGE_F32, j->master, storage, vol, nvox);
        else e = convert_any(b->stream, j->d.storage, j->vol, storage, vol, nvox);
    }
    if (e == hipSuccess) e = convert_any(b->stream, ps_old, j->exps, ps_new, exps, npx);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (vol) (void)hipFree(vol);
        if (exps) (void)hipFree(exps);
        if (diff) (void)hipFree(diff);
        if (pairs) (void)hipFree(pairs);
        return bfail(TJ_EDEVICE, "tj_batch_convert: %s", hipGetErrorString(e));
    }
    (void)hipFree(j->vol);
    (void)hipFree((void*)j->exps);
    j->vol = vol;
    j->exps = exps;
    if (diff) {
        (void)hipFree(j->diff);
        j->diff = diff;
    }
    if (pairs) j->pairs = pairs;
    if (!half_new && j->pairs) {
        (void)hipFree(j->pairs);
        j->pairs = nullptr;
    }
    // a 16-bit target keeps the master when the job asked for one
    if (j->master && !(storage == TJ_STORAGE_BF16 || storage == TJ_STORAGE_F16)) {
        (void)hipFree(j->master);
        j->master = nullptr;
    }
    b->used = b->used - j->bytes + need;
    j->bytes = need;
    j->d.storage = storage;
    drop_plans(b, id);
    return TJ_OK;
}

int tj_batch_reset_volume(tj_batch* b, int id)
{
    if (!b) return bfail(TJ_EINVAL, "null batch");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    const size_t nvox = (size_t)j->d.vol_w * j->d.vol_h * j->d.vol_d;
    TB_HIP(hipMemsetAsync(j->vol, 0, nvox * storage_bytes(j->d.storage), b->stream));
    if (j->master) TB_HIP(hipMemsetAsync(j->master, 0, nvox * 4, b->stream));
    TB_HIP(hipMemsetAsync(j->iter, 0, sizeof(int), b->stream));
    TB_HIP(hipStreamSynchronize(b->stream));
    return TJ_OK;
}

int tj_batch_run(tj_batch* b, const int* ids, int nids, int iterations, int flags)
{
    if (!b || !ids || nids <= 0 || iterations < 0) return bfail(TJ_EINVAL, "tj_batch_run: bad argument");
    TB_HIP(hipSetDevice(b->device));
    std::vector<int> v(ids, ids + nids);
    {
        std::vector<int> s = v;
        std::sort(s.begin(), s.end());
        if (std::adjacent_find(s.begin(), s.end()) != s.end()) return bfail(TJ_EINVAL, "tj_batch_run: repeated job");
    }
    if (iterations == 0) return TJ_OK;
    const bool graph = (flags & TJ_RUN_GRAPH) != 0;
    const bool reuse = plan_matches(b, v);
    if (!reuse) {
        if (int rc = build_plan(b, v)) return rc;
    }
    RunPlan& p = *b->cur;
    TB_HIP(hipEventRecord(b->ev0, b->stream));
    if (graph) {
        if (!(reuse && p.exec && p.iterations == iterations)) {
            if (p.exec) (void)hipGraphExecDestroy(p.exec);
            if (p.graph) (void)hipGraphDestroy(p.graph);
            p.exec = nullptr;
            p.graph = nullptr;
            TB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_iterations(b, iterations);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(TJ_EDEVICE, "tj_batch_run: graph capture failed: %s",
                             hipGetErrorString(e != hipSuccess ? e : e2));
            }
            p.graph = g;
            TB_HIP(hipGraphInstantiate(&p.exec, g, nullptr, nullptr, 0));
            p.iterations = iterations;
        }
        TB_HIP(hipGraphLaunch(p.exec, b->stream));
    } else {
        TB_HIP(enqueue_iterations(b, iterations));
    }
    TB_HIP(hipEventRecord(b->ev1, b->stream));
    TB_HIP(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    TB_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    return TJ_OK;
}

double tj_batch_last_run_ms(const tj_batch* b) { return b ? b->last_ms : 0.0; }

int tj_batch_errors(tj_batch* b, int id, double* err, int cap)
{
    if (!b) return bfail(TJ_EINVAL, "null batch");
    BJob* j = find_job(b, id);
    if (!j) return bfail(TJ_EINVAL, "unknown job %d", id);
    TB_HIP(hipSetDevice(b->device));
    int it = 0;
    TB_HIP(hipMemcpy(&it, j->iter, sizeof(int), hipMemcpyDeviceToHost));
    if (err && cap > 0) {
        const int hc = std::max(j->d.max_iterations, 1);
        std::vector<double> h(hc);
        TB_HIP(hipMemcpy(h.data(), j->err_hist, (size_t)hc * 8, hipMemcpyDeviceToHost));
        // the last min(it, hc, cap) iterations, oldest first
        const int m = std::min(std::min(it, hc), cap);
        for (int k = 0; k < m; ++k) err[k] = h[(size_t)(it - m + k) % hc];
    }
    return it;
}

int tj_batch_set_tuning(tj_batch* b, const tj_batch_tuning* t)
{
    if (!b || !t) return bfail(TJ_EINVAL, "null argument");
    if (t->fp_rays < 1 || t->fp_lds_kb < 16 || t->fp_lds_kb > 152 || t->bp_rows < 1 || t->bp_zchunk < 1 ||
        t->bp_group < 1 || t->blocks_per_cu < 1 || t->blocks_per_cu > 16 || t->fp_packed < 0 || t->fp_packed > 2 ||
        t->fp_cols < 0 || t->fp_cols > 2 || t->fp_shared < 0 || t->fp_shared > 1)
        return bfail(TJ_EINVAL, "tj_batch_set_tuning: value out of range");
    b->tune = *t;
    return TJ_OK;
}

int tj_batch_get_tuning(const tj_batch* b, tj_batch_tuning* t)
{
    if (!b || !t) return bfail(TJ_EINVAL, "null argument");
    *t = b->tune;
    return TJ_OK;
}

// Times one launch kind of the current plan (forward = 0, back = 2) over
// `reps` repetitions; volumes are saved before and restored after when the
// backprojection is timed.
static int time_kind(tj_batch* b, int kind, int reps, double* ms_out)
{
    RunPlan& p = *b->cur;
    // (device buffer, saved copy, bytes) of every volume the backprojection writes
    struct Saved {
        void* buf;
        void* copy;
        size_t bytes;
    };
    std::vector<Saved> saved;
    auto release = [&]() {
        for (auto& q : saved) (void)hipFree(q.copy);
    };
    if (kind == 2) {
        for (int id : p.ids) {
            BJob* j = find_job(b, id);
            const size_t nvox = (size_t)j->d.vol_w * j->d.vol_h * j->d.vol_d;
            std::pair<void*, size_t> bufs[2] = {{j->vol, nvox * storage_bytes(j->d.storage)},
                                                {j->master, j->master ? nvox * 4 : 0}};
            for (auto& bf : bufs) {
                if (!bf.first) continue;
                void* c = nullptr;
                if (hipMalloc(&c, bf.second) != hipSuccess) {
                    (void)hipGetLastError();
                    release();
                    return bfail(TJ_ENOMEM, "autotune: no room to save the volumes");
                }
                saved.push_back({bf.first, c, bf.second});
                if (hipMemcpyAsync(c, bf.first, bf.second, hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
                    release();
                    return bfail(TJ_EDEVICE, "autotune: cannot save a volume");
                }
            }
        }
    }
    hipError_t e = hipSuccess;
    for (int r = 0; r <= reps && e == hipSuccess; ++r) {
        if (r == 1) e = hipEventRecord(b->ev0, b->stream);
        for (const RunPlan::L& l : p.launches)
            if (l.kind == kind && e == hipSuccess) e = enqueue_launch(b, l);
    }
    if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
    for (auto& q : saved) (void)hipMemcpyAsync(q.buf, q.copy, q.bytes, hipMemcpyDeviceToDevice, b->stream);
    (void)hipStreamSynchronize(b->stream);
    release();
    if (e != hipSuccess) return bfail(TJ_EDEVICE, "autotune: %s", hipGetErrorString(e));
    *ms_out = ms / std::max(reps, 1);
    return TJ_OK;
}

int tj_batch_autotune(tj_batch* b, const int* ids, int nids, tj_batch_tuning* best, double* best_ms)
{
    if (!b || !ids || nids <= 0) return bfail(TJ_EINVAL, "tj_batch_autotune: bad argument");
    TB_HIP(hipSetDevice(b->device));
    std::vector<int> v(ids, ids + nids);
    const tj_batch_tuning start = b->tune;
    // forward: rays per tile x LDS per tile x persistent blocks per CU
    const int fr[] = {2, 4}, fl[] = {32, 64, 96, 128}, bpc[] = {2, 4, 8};
    bool any_half = false;
    for (int id : v) {
        BJob* j = find_job(b, id);
        if (j && (j->d.storage == TJ_STORAGE_BF16 || j->d.storage == TJ_STORAGE_F16)) any_half = true;
    }
    tj_batch_tuning cur = start, bestf = start;
    double tf = 1e300;
    // the double-buffered packed kernel has one shape (its rays, LDS and
    // one block per CU are fixed)
    // k_job_fq (shared, double-buffered) likewise has one shape
    for (int pk = any_half ? 0 : 1; pk <= (any_half ? 2 : 1); ++pk)
    for (int sh = 0; sh <= 1; ++sh)
    for (int cols = 1; cols <= ((any_half && pk == 1) ? 2 : 1); ++cols)
    for (int r : fr)
        if ((pk != 2 || r == kFpqRays) && (!sh || r == kFqRays))
        for (int l : fl)
            if ((pk != 2 && !sh) || l == fl[0])
            for (int c : bpc)
                if ((pk != 2 && !sh) || c == bpc[0]) {
                cur.fp_packed = any_half ? pk : start.fp_packed;
                cur.fp_cols = (any_half && pk) ? cols : start.fp_cols;
                cur.fp_shared = sh;
                cur.fp_rays = r;
                cur.fp_lds_kb = l;
                cur.blocks_per_cu = c;
                b->tune = cur;
                if (int rc = build_plan(b, v)) return rc;
                double ms = 0;
                if (int rc = time_kind(b, 0, 2, &ms)) return rc;
                if (ms < tf) {
                    tf = ms;
                    bestf = cur;
                }
            }
    // backprojection: (rows, slices) x persistent blocks per CU, forward fixed
    const int shapes[][2] = {{4, 8}, {2, 8}, {4, 4}, {2, 16}, {4, 16}, {8, 8}, {8, 16}};
    cur = bestf;
    tj_batch_tuning bestb = bestf;
    double tb = 1e300;
    for (auto& sh : shapes)
        for (int c : bpc) {
            cur.bp_rows = sh[0];
            cur.bp_zchunk = sh[1];
            cur.bp_group = (sh[1] >= 16 || sh[0] >= 8) ? 2 : 4;
            cur.blocks_per_cu = c;
            b->tune = cur;
            if (int rc = build_plan(b, v)) return rc;
            double ms = 0;
            if (int rc = time_kind(b, 2, 2, &ms)) return rc;
            if (ms < tb) {
                tb = ms;
                bestb = cur;
            }
        }
    // one persistent block count serves both kernels: keep the one that is
    // better for the sum
    tj_batch_tuning fin = bestb;
    fin.fp_rays = bestf.fp_rays;
    fin.fp_lds_kb = bestf.fp_lds_kb;
    fin.fp_packed = bestf.fp_packed;
    fin.fp_cols = bestf.fp_cols;
    fin.fp_shared = bestf.fp_shared;
    double tot_best = 1e300;
    tj_batch_tuning pick = fin;
    for (int c : {bestf.blocks_per_cu, bestb.blocks_per_cu}) {
        fin.blocks_per_cu = c;
        b->tune = fin;
        if (int rc = build_plan(b, v)) return rc;
        double a = 0, c2 = 0;
        if (int rc = time_kind(b, 0, 2, &a)) return rc;
        if (int rc = time_kind(b, 2, 2, &c2)) return rc;
        if (a + c2 < tot_best) {
            tot_best = a + c2;
            pick = fin;
        }
    }
    b->tune = pick;
    drop_plans_of_set(b, v);
    if (best) *best = pick;
    if (best_ms) *best_ms = tot_best;
    return TJ_OK;
}

}  // extern "C"
