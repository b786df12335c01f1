// tj_kernels.hip -- CDNA4 (gfx950) kernels of the TomoJ projector path.
//
// Arithmetic follows TomoJ's CPU projector operation by operation so that the
// FP32 results are the ones the Java code produces: positions in FP64, samples
// and accumulators in FP32, casts at the same points, and no fused multiply-add
// (this file is built with -ffp-contract=off).  Reference methods are cited per
// kernel (paths relative to src/TomoJ_Base/src/main/java/fr/curie/tomoj/).
#include "tj_device.h"
#include <math.h>
#include <type_traits>

// LDS (address space 3) pointers: some kernels keep LDS byte addresses as
// 32-bit values in registers and turn them into such pointers where they read
// them; the host pass, which sees 64-bit pointers, would warn about each cast.
#define TJ_LDS __attribute__((address_space(3)))
#pragma clang diagnostic ignored "-Wint-to-pointer-cast"

namespace tj {

// ---------------------------------------------------------------------------
// Block reductions with a fixed combination order (deterministic results).
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ inline T block_reduce(T v, T* sh, Op op)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] = op(sh[t], sh[t + off]);
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

struct AddD { __device__ double operator()(double a, double b) const { return a + b; } };
struct MaxF { __device__ float operator()(float a, float b) const { return a > b ? a : b; } };
struct MaxD { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
struct OrU { __device__ unsigned operator()(unsigned a, unsigned b) const { return a | b; } };
struct MaxU { __device__ unsigned operator()(unsigned a, unsigned b) const { return a > b ? a : b; } };
struct MinD { __device__ double operator()(double a, double b) const { return a < b ? a : b; } };

// jcentered of detector row j: VoxelProjector3D.project splits the rows from
// `start` into tasks of two (:350-352) and projectJosephPartial sets
// jcentered = (starty - 1.0 - pc1)/scale, then adds 1/scale per row (:711-724).
__device__ inline double jcentered(const TjGeomDev& g, int start, int j)
{
    const int sy = start + ((j - start) & ~1);
    const double jinc = 1.0 / g.scale;
    double jc = (sy - 1.0 - g.pc1) / g.scale;
    jc += jinc;
    if (j != sy) jc += jinc;
    return jc;
}

// TomoReconstruction2.getPixel(double x, double y, int z) (TomoReconstruction2.java:154-190)
// on the [y][z][x] device layout.
__device__ inline float getpixel(const TjGeomDev& g, const float* __restrict__ vol, double x,
                                 double y, int z)
{
    const int xb = (x < 0) ? 0 : (int)x;
    const int yb = (y < 0) ? 0 : (int)y;
    const size_t plane = (size_t)g.vd * g.vw;
    const float* p = vol + (size_t)yb * plane + (size_t)z * g.vw + xb;
    if (xb == g.vw - 1 || yb == g.vh - 1 || z == g.vd - 1) return p[0];
    const float xf = (x < 0) ? 0.0f : (float)x - (float)xb;
    const float yf = (y < 0) ? 0.0f : (float)y - (float)yb;
    const float v00 = p[0], v10 = p[1];
    const float v01 = p[plane], v11 = p[plane + 1];
    const float lower = v00 + xf * (v10 - v00);
    return lower + yf * ((v01 + xf * (v11 - v01)) - lower);
}

// ---------------------------------------------------------------------------
// Forward projection, any geometry: one thread per ray (slot, i, j).
// VoxelProjector3D.projectJosephPartial (VoxelProjector3D.java:671-821).
// Rows below `start_row` are left at 0 (TomoJ's fresh FloatProcessor) (:343-347).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TJ_BLOCK) void k_fp_generic(
    TjGeomDev g, const float* __restrict__ vol, const TjProj* __restrict__ pp,
    const int* __restrict__ slot_proj, int start_row, float* __restrict__ proj,
    float* __restrict__ norm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int s = blockIdx.z;
    if (i >= g.pw) return;
    const size_t o = ((size_t)s * g.ph + j) * g.pw + i;
    if (j < start_row) {
        proj[o] = 0.0f;
        norm[o] = 0.0f;
        return;
    }
    const TjProj& P = pp[slot_proj[s]];
    const double jc = jcentered(g, start_row, j);
    const double ic = (i - g.pc0) / g.scale;
    float ray_sum = 0.0f, norm_sum = 0.0f;
    const double ijpart = jc - P.icf * ic;
    double oldy = (ijpart - P.zcf * (-1 - g.vcz)) / P.divy;
    double oldx = ic - P.yfx * oldy - P.zfx * (-1 - g.vcz);
    oldx /= P.divx;
    bool entered = false;
    for (int z = 0; z < g.vd; ++z) {
        const double xc = oldx + P.dirx;
        const double yc = oldy + P.diry;
        const double x = xc + g.vcx;
        const double y = yc + g.vcy;
        if (x + TJ_ACCURACY >= 0 && x < g.bxmax && y + TJ_ACCURACY >= 0 && y < g.bymax) {
            ray_sum += getpixel(g, vol, x, y, z);
            norm_sum = (float)((double)norm_sum + P.norminc);
            entered = true;
        } else if (entered) {
            norm_sum = (float)((double)norm_sum + P.norminc);
            break;
        }
        oldx = xc;
        oldy = yc;
    }
    if (norm_sum > 0) {
        proj[o] = (float)((double)ray_sum * P.norminc);
        norm[o] = norm_sum;
    } else {
        proj[o] = __builtin_nanf("");
        norm[o] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Forward projection, single tilt axis (e01 == e10 == e12 == 0): every ray of a
// detector column follows the same x(z) sequence and keeps a constant y, so one
// thread marches one column (slot, i) and accumulates R detector rows with it.
// Same arithmetic as k_fp_generic, sample for sample.  pflag[y] != 0 marks a
// volume plane holding non-finite values: rows touching such a plane (or with a
// fractional y) take the full 4-tap getPixel path.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(TJ_BLOCK) void k_fp_axis(
    TjGeomDev g, const float* __restrict__ vol, const TjProj* __restrict__ pp,
    const int* __restrict__ slot_proj, int start_row, const uint8_t* __restrict__ pflag,
    float* __restrict__ proj, float* __restrict__ norm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j0 = blockIdx.y * R;
    const int s = blockIdx.z;
    if (i >= g.pw) return;
    const TjProj& P = pp[slot_proj[s]];
    const double ic = (i - g.pc0) / g.scale;
    const size_t plane = (size_t)g.vd * g.vw;

    const float* rowp[R];
    float yf[R];
    bool ok[R], nearY[R], full[R];
    double oldy0 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        ok[r] = false;
        nearY[r] = false;
        full[r] = false;
        yf[r] = 0.0f;
        rowp[r] = vol;
        if (j < g.ph && j >= start_row) {
            const double jc = jcentered(g, start_row, j);
            const double ijpart = jc - P.icf * ic;
            const double oldy = (ijpart - P.zcf * (-1 - g.vcz)) / P.divy;
            oldy0 = oldy;
            const double y = (oldy + P.diry) + g.vcy;
            if (y + TJ_ACCURACY >= 0 && y < g.bymax) {
                ok[r] = true;
                const int yb = (y < 0) ? 0 : (int)y;
                const float yfr = (y < 0) ? 0.0f : (float)y - (float)yb;
                rowp[r] = vol + (size_t)yb * plane;
                nearY[r] = (yb == g.vh - 1);
                full[r] = !nearY[r] && (yfr != 0.0f || pflag[yb] != 0 || pflag[yb + 1] != 0);
                yf[r] = yfr;
            }
        }
    }
    // yfx == 0 here, so oldx does not depend on the row's oldy.
    double xc = ic - P.yfx * oldy0 - P.zfx * (-1 - g.vcz);
    xc /= P.divx;

    float rs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rs[r] = 0.0f;
    float ns = 0.0f;
    bool entered = false;
    for (int z = 0; z < g.vd; ++z) {
        xc = xc + P.dirx;
        const double x = xc + g.vcx;
        if (x + TJ_ACCURACY >= 0 && x < g.bxmax) {
            const int xb = (x < 0) ? 0 : (int)x;
            const float xf = (x < 0) ? 0.0f : (float)x - (float)xb;
            const bool nearXZ = (xb == g.vw - 1) || (z == g.vd - 1);
            const size_t zo = (size_t)z * g.vw + xb;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (ok[r]) {
                    const float* p = rowp[r] + zo;
                    float v;
                    if (nearXZ || nearY[r]) {
                        v = p[0];
                    } else {
                        const float v00 = p[0], v10 = p[1];
                        const float lower = v00 + xf * (v10 - v00);
                        if (full[r]) {
                            const float v01 = p[plane], v11 = p[plane + 1];
                            v = lower + yf[r] * ((v01 + xf * (v11 - v01)) - lower);
                        } else {
                            v = lower;
                        }
                    }
                    rs[r] += v;
                }
            }
            ns = (float)((double)ns + P.norminc);
            entered = true;
        } else if (entered) {
            ns = (float)((double)ns + P.norminc);
            break;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j >= g.ph) break;
        const size_t o = ((size_t)s * g.ph + j) * g.pw + i;
        if (j < start_row) {
            proj[o] = 0.0f;
            norm[o] = 0.0f;
        } else if (ok[r] && ns > 0) {
            proj[o] = (float)((double)rs[r] * P.norminc);
            norm[o] = ns;
        } else {
            proj[o] = __builtin_nanf("");
            norm[o] = 0.0f;
        }
    }
}

// Long-object compensation: the norm image becomes its maximum
// (VoxelProjector3D.java:799-801, :813-819).  One block per slot.
__global__ __launch_bounds__(TJ_BLOCK) void k_norm_max_fill(int npx, float* __restrict__ norm)
{
    __shared__ float sh[TJ_BLOCK];
    float* n = norm + (size_t)blockIdx.x * npx;
    float m = 0.0f;
    for (int k = threadIdx.x; k < npx; k += blockDim.x) m = fmaxf(m, n[k]);
    m = block_reduce(m, sh, MaxF());
    for (int k = threadIdx.x; k < npx; k += blockDim.x) n[k] = m;
}

// ---------------------------------------------------------------------------
// Residual: VoxelProjector3D.difference (VoxelProjector3D.java:414-433).  One
// block per (row j, slot s).  diff = (float)((double)((exp - th) / norm) * factor),
// left 0 where exp or th is NaN; row_err[s][j] = sum of value^2 over the row in
// FP64; rowflag[s][j] = 1 if the row of diff holds a non-finite value.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TJ_BLOCK) void k_residual(
    int pw, int ph, const float* __restrict__ exps, const int* __restrict__ exp_index,
    const float* __restrict__ proj, const float* __restrict__ norm, double factor,
    float* __restrict__ diff, double* __restrict__ row_err, uint8_t* __restrict__ rowflag)
{
    __shared__ double shd[TJ_BLOCK];
    __shared__ unsigned shu[TJ_BLOCK];
    const int j = blockIdx.x, s = blockIdx.y;
    const size_t npx = (size_t)pw * ph;
    const float* e = exps + (size_t)exp_index[s] * npx + (size_t)j * pw;
    const size_t o = (size_t)s * npx + (size_t)j * pw;
    double acc = 0.0;
    unsigned bad = 0;
    for (int i = threadIdx.x; i < pw; i += blockDim.x) {
        const float ex = e[i], th = proj[o + i], n = norm[o + i];
        float d = 0.0f;
        if (!isnan(ex) && !isnan(th)) {
            const double value = (double)((ex - th) / n) * factor;
            d = (float)value;
            acc += value * value;
        }
        diff[o + i] = d;
        bad |= isfinite(d) ? 0u : 1u;
    }
    acc = block_reduce(acc, shd, AddD());
    bad = block_reduce(bad, shu, OrU());
    if (threadIdx.x == 0) {
        row_err[(size_t)s * ph + j] = acc;
        rowflag[(size_t)s * ph + j] = (uint8_t)bad;
    }
}

// Sum of the per-row errors of each slot in a fixed order; one block per slot.
__global__ __launch_bounds__(TJ_BLOCK) void k_err_sum(int ph, const double* __restrict__ row_err,
                                                      double* __restrict__ err)
{
    __shared__ double sh[TJ_BLOCK];
    double acc = 0.0;
    for (int j = threadIdx.x; j < ph; j += blockDim.x) acc += row_err[(size_t)blockIdx.x * ph + j];
    acc = block_reduce(acc, sh, AddD());
    if (threadIdx.x == 0) err[blockIdx.x] = acc;
}

// Non-finite flag per image row for images that did not come from k_residual
// (weighted or original projections fed to the backprojection).
__global__ __launch_bounds__(TJ_BLOCK) void k_row_flags(int pw, int ph, const float* __restrict__ images,
                                                        const int* __restrict__ image_index,
                                                        uint8_t* __restrict__ rowflag)
{
    __shared__ unsigned sh[TJ_BLOCK];
    const int j = blockIdx.x, s = blockIdx.y;
    const float* p = images + (size_t)image_index[s] * pw * ph + (size_t)j * pw;
    unsigned bad = 0;
    for (int i = threadIdx.x; i < pw; i += blockDim.x) bad |= isfinite(p[i]) ? 0u : 1u;
    bad = block_reduce(bad, sh, OrU());
    if (threadIdx.x == 0) rowflag[(size_t)s * ph + j] = (uint8_t)bad;
}

// ---------------------------------------------------------------------------
// Backprojection.  VoxelProjector3D.backProjectPartial (VoxelProjector3D.java:221-308)
// with ImageJ's getInterpolatedValue (bilinear in double, neighbours clamped to
// the last column/row).  Each voxel takes the added images in order, with the
// NaN/Inf scrub and the positivity clamp after every image, exactly as the
// Java loop does one projection at a time.  Each voxel is owned by one thread:
// no atomics, results independent of scheduling.
// ---------------------------------------------------------------------------

// Row table for the single-axis kernel: yy = ((x*e10 + y*e11) + z*e12)*scale + ciy
// reduces to (y*e11)*scale + ciy when e10 == e12 == 0.
__device__ __forceinline__ void bp_row(const TjGeomDev& g, const TjProj* __restrict__ pp,
                                       const int* __restrict__ slot_proj, int nslot,
                                       const uint8_t* __restrict__ rowflag, TjBpRow* __restrict__ rows,
                                       int* __restrict__ slow, int j, int s)
{
    if (j >= g.vh || s >= nslot) return;
    const TjProj& Q = pp[slot_proj[s]];
    const double y = j - g.vcy;
    const double yy = (y * Q.e[4]) * g.scale + g.pc3;
    TjBpRow r;
    r.valid = (yy >= 0 && yy < g.ph) ? 1 : 0;
    r.yb = 0;
    r.y1 = 0;
    r.yf = 0.0;
    r.shortcut = 1;
    r.finite = 1;
    r.pad = 0;
    if (r.valid) {
        r.yb = (int)yy;
        r.y1 = min(r.yb + 1, g.ph - 1);
        r.yf = yy - (double)r.yb;
        r.shortcut = (r.y1 == r.yb) || (r.yf == 0.0 && rowflag[(size_t)s * g.ph + r.y1] == 0);
        r.finite = rowflag[(size_t)s * g.ph + r.yb] == 0;
        // a row the fast backprojection cannot take (fractional yy, or a
        // non-finite value in its image rows): the general kernel runs
        if (slow && !(r.shortcut && r.finite)) atomicOr(slow, 1);
    }
    rows[(size_t)s * g.vh + j] = r;
}

__global__ void k_bp_rows(TjGeomDev g, const TjProj* __restrict__ pp, const int* __restrict__ slot_proj,
                          int nslot, const uint8_t* __restrict__ rowflag, TjBpRow* __restrict__ rows,
                          int* __restrict__ slow)
{
    bp_row(g, pp, slot_proj, nslot, rowflag, rows, slow, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

template <bool POS>
__device__ inline void bp_accumulate(float& acc, float value)
{
    if (!isnan(value)) {
        float a = acc + value;
        if (isnan(a) || isinf(a)) a = 0.0f;
        if (POS && a < 0) a = 0.0f;
        acc = a;
    }
}

// Single-axis backprojection: thread = voxel column x=i over R volume rows and
// a chunk of z; xx depends on (slot, i, k) only and yy on (slot, j) only.
template <int R, bool POS, bool SCALE1>
__global__ __launch_bounds__(TJ_BLOCK) void k_bp_axis(
    TjGeomDev g, float* __restrict__ vol, const TjProj* __restrict__ pp,
    const int* __restrict__ slot_proj, int nslot, const float* __restrict__ images,
    const int* __restrict__ image_index, const TjBpRow* __restrict__ rows, int start_y, int end_y,
    int kchunk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.vw) return;
    const int j0 = start_y + blockIdx.y * R;
    const int k0 = blockIdx.z * kchunk;
    const int k1 = min(k0 + kchunk, g.vd);
    const int nr = min(R, end_y - j0);
    const double x = i - g.vcx;
    const size_t plane = (size_t)g.vd * g.vw;
    const size_t npx = (size_t)g.pw * g.ph;
    for (int k = k0; k < k1; ++k) {
        const double z = k - g.vcz;
        float acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[r] = (r < nr) ? vol[(size_t)(j0 + r) * plane + (size_t)k * g.vw + i] : 0.0f;
        for (int s = 0; s < nslot; ++s) {
            const TjProj& Q = pp[slot_proj[s]];
            // ((x*e00 + y*e01) + z*e02): y*e01 is a signed zero here, which changes
            // no value the comparisons and interpolation below can see.
            double xx = x * Q.e[0] + z * Q.e[2];
            if (!SCALE1) xx = xx * g.scale;
            xx = xx + g.pc2;
            if (!(xx >= 0 && xx < g.pw)) continue;
            const int xb = (int)xx;
            const double xf = xx - (double)xb;
            const int x1 = min(xb + 1, g.pw - 1);
            const float* img = images + (size_t)image_index[s] * npx;
            const TjBpRow* rw = rows + (size_t)s * g.vh + j0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r >= nr) break;
                const TjBpRow row = rw[r];
                if (!row.valid) continue;
                const float* lo = img + (size_t)row.yb * g.pw;
                const double ll = lo[xb], lr = lo[x1];
                const double lower = ll + xf * (lr - ll);
                float value;
                if (row.shortcut) {
                    // lower + 0*(upper - lower): lower when finite, NaN otherwise
                    value = isfinite(lower) ? (float)lower : __builtin_nanf("");
                } else {
                    const float* up = img + (size_t)row.y1 * g.pw;
                    const double ul = up[xb], ur = up[x1];
                    const double upper = ul + xf * (ur - ul);
                    value = (float)(lower + row.yf * (upper - lower));
                }
                bp_accumulate<POS>(acc[r], value);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < nr) vol[(size_t)(j0 + r) * plane + (size_t)k * g.vw + i] = acc[r];
    }
}

// Any-geometry backprojection: one thread per voxel (i, j, k).
template <bool POS>
__global__ __launch_bounds__(TJ_BLOCK) void k_bp_generic(
    TjGeomDev g, float* __restrict__ vol, const TjProj* __restrict__ pp,
    const int* __restrict__ slot_proj, int nslot, const float* __restrict__ images,
    const int* __restrict__ image_index, int start_y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = start_y + blockIdx.y;
    const int k = blockIdx.z;
    if (i >= g.vw) return;
    const double x = i - g.vcx, y = j - g.vcy, z = k - g.vcz;
    const size_t npx = (size_t)g.pw * g.ph;
    float* v = vol + ((size_t)j * g.vd + k) * g.vw + i;
    float acc = *v;
    for (int s = 0; s < nslot; ++s) {
        const TjProj& Q = pp[slot_proj[s]];
        const double ztx = z * Q.e[2], zty = z * Q.e[5];
        const double ytx = y * Q.e[1], yty = y * Q.e[4];
        const double xx = (x * Q.e[0] + ytx + ztx) * g.scale + g.pc2;
        const double yy = (x * Q.e[3] + yty + zty) * g.scale + g.pc3;
        if (xx >= 0 && xx < g.pw && yy >= 0 && yy < g.ph) {
            const float* img = images + (size_t)image_index[s] * npx;
            const int xb = (int)xx, yb = (int)yy;
            const int x1 = min(xb + 1, g.pw - 1), y1 = min(yb + 1, g.ph - 1);
            const double xf = xx - (double)xb, yf = yy - (double)yb;
            const double ll = img[(size_t)yb * g.pw + xb], lr = img[(size_t)yb * g.pw + x1];
            const double ul = img[(size_t)y1 * g.pw + xb], ur = img[(size_t)y1 * g.pw + x1];
            const double upper = ul + xf * (ur - ul);
            const double lower = ll + xf * (lr - ll);
            bp_accumulate<POS>(acc, (float)(lower + yf * (upper - lower)));
        }
    }
    *v = acc;
}

// ---------------------------------------------------------------------------
// LDS-staged single-axis kernels (the fast path for tilt series around a
// vertical axis with scale 1).  Same arithmetic as the kernels above.
// ---------------------------------------------------------------------------

// FP64 pair image of the images a backprojection adds: pairs[s][y][x] =
// ((double)ll, (double)lr - (double)ll) with ll = image(s)[y][x], lr its right
// neighbour (the last column repeated) -- the two operands of ImageJ's
// interpolation along x, in the cells' order, so the backprojection stages
// them with plain buffer-to-LDS copies.
// Pair images of slot s: blocks b0 = 0 .. nb - 1 of a grid-stride loop.
__device__ __forceinline__ void bp_pairs(const float* __restrict__ images, const int* __restrict__ image_index,
                                         int pw, int ph, double2* __restrict__ pairs, int s, int b0, int nb)
{
    const float* img = images + (size_t)image_index[s] * pw * ph;
    double2* out = pairs + (size_t)s * pw * ph;
    const size_t npx = (size_t)pw * ph;
    for (size_t k = (size_t)b0 * TJ_BLOCK + threadIdx.x; k < npx; k += (size_t)nb * TJ_BLOCK) {
        const int x = (int)(k % pw);
        const double ll = img[k], lr = img[x + 1 < pw ? k + 1 : k];
        out[k] = make_double2(ll, lr - ll);
    }
}

// The backprojection's preparation in one launch: blocks [0, nrb) of each slot
// (grid y) build its row table (k_bp_rows), the others its pair images
// (k_bp_pairs).  The two parts are independent.
__global__ __launch_bounds__(TJ_BLOCK) void k_bp_prep(TjGeomDev g, const TjProj* __restrict__ pp,
                                                      const int* __restrict__ slot_proj, int nslot,
                                                      const uint8_t* __restrict__ rowflag,
                                                      TjBpRow* __restrict__ rows, int* __restrict__ slow, int nrb,
                                                      const float* __restrict__ images,
                                                      const int* __restrict__ image_index,
                                                      double2* __restrict__ pairs)
{
    if ((int)blockIdx.x < nrb)
        bp_row(g, pp, slot_proj, nslot, rowflag, rows, slow, blockIdx.x * TJ_BLOCK + threadIdx.x, blockIdx.y);
    else
        bp_pairs(images, image_index, g.pw, g.ph, pairs, blockIdx.y, blockIdx.x - nrb, gridDim.x - nrb);
}

// Backprojection: a 256-thread block owns 256 voxel columns x, R volume rows and
// KC slices, and keeps their KC*R sums in registers.  For each group of G added
// images it stages, per (image, row), the detector columns its voxels can reach
// as FP64 pairs (ll, lr - ll): exactly the two operands of ImageJ's
// interpolation along x, so a voxel update is one LDS read, one multiply and one
// add before the float cast.  Images are taken in order for every voxel.  The
// update is branch-free: lanes whose projection falls outside the image read a
// fixed LDS cell and keep their sum.  Rows whose y-interpolation does not reduce
// to the lower row (fractional yy, or a non-finite value in the next image row)
// take the exact 4-tap path with the next row read from global memory.
template <int R, int KC, int G, bool POS>
__global__ __launch_bounds__(TJ_BLOCK) void k_bp_lds(
    TjGeomDev g, float* __restrict__ vol, const TjProj* __restrict__ pp,
    const int* __restrict__ slot_proj, int nslot, const float* __restrict__ images,
    const int* __restrict__ image_index, const TjBpRow* __restrict__ rows, int start_y, int end_y,
    int finite_volume, const double2* __restrict__ pairs, const int* __restrict__ run_if)
{
    if (run_if && *run_if == 0) return;  // the fast kernel takes this launch (uniform)
    constexpr int XW = TJ_BLOCK + KC + 8;  // reachable columns + a zero cell at XW-1
    constexpr int NB = (XW + 63) / 64;     // 64-cell copies per staged row
    __shared__ double2 lds[G][R][XW];
    // per-slot constants of the current group, staged once per pass
    __shared__ double s_e0[G], s_e2[G];
    __shared__ int s_x0[G], s_mode[G], s_img[G];
    __shared__ double s_ryf[G][R];
    __shared__ int s_rflag[G][R];  // bit0 valid, bit1 shortcut
    __shared__ int s_ryb[G][R], s_ry1[G][R];
    const int tid = threadIdx.x;
    // XCD-aware block order (1-D grid of nbx * nby * nbz blocks): blocks are
    // dealt to the 8 XCDs round-robin, so block b runs on XCD b % 8; all blocks
    // of one row group (the ones that stage the same image rows) are put on one
    // XCD, consecutive there, so its L2 serves their copies.  With a row-group
    // count that is not a multiple of 8 the order is plain x, y, z.
    const int nbx = (g.vw + TJ_BLOCK - 1) / TJ_BLOCK, nby = (end_y - start_y + R - 1) / R;
    const int nbz = (g.vd + KC - 1) / KC;
    int bx, by, bz;
    {
        const int b = blockIdx.x;
        if ((nby & 7) == 0) {
            const int per = nbx * nbz, k = b >> 3;
            by = (b & 7) + 8 * (k / per);
            const int rr = k - (k / per) * per;
            bz = rr / nbx;
            bx = rr - bz * nbx;
        } else {
            bx = b % nbx;
            by = (b / nbx) % nby;
            bz = b / (nbx * nby);
        }
    }
    const int i0 = bx * TJ_BLOCK;
    const int i = i0 + tid;
    const int j0 = start_y + by * R;
    const int nr = min(R, end_y - j0);
    const int k0 = bz * KC;
    const int nk = min(KC, g.vd - k0);
    const int ilast = min(i0 + TJ_BLOCK - 1, g.vw - 1);
    const bool active = i < g.vw;
    const double x = (active ? i : ilast) - g.vcx;
    const size_t plane = (size_t)g.vd * g.vw;
    const size_t npx = (size_t)g.pw * g.ph;
    const double pwd = g.pw;

    float acc[KC][R];
#pragma unroll
    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[kk][r] = (active && kk < nk && r < nr)
                             ? vol[(size_t)(j0 + r) * plane + (size_t)(k0 + kk) * g.vw + i]
                             : 0.0f;

    for (int s0 = 0; s0 < nslot; s0 += G) {
        const int gn = min(G, nslot - s0);
        __syncthreads();
        if (tid < gn) {
            // smallest xx over the block: xx is monotone in i and in k, so it is
            // at a corner; the same expression as the per-voxel xx below.
            const TjProj& Q = pp[slot_proj[s0 + tid]];
            const double xa = (i0 - g.vcx) * Q.e[0], xb = (ilast - g.vcx) * Q.e[0];
            const double za = (k0 - g.vcz) * Q.e[2], zb = (k0 + nk - 1 - g.vcz) * Q.e[2];
            const double m = fmin(fmin((xa + za) + g.pc2, (xa + zb) + g.pc2),
                                  fmin((xb + za) + g.pc2, (xb + zb) + g.pc2));
            s_x0[tid] = (m > 0.0) ? (int)m : 0;
            s_e0[tid] = Q.e[0];
            s_e2[tid] = Q.e[2];
            s_img[tid] = image_index[s0 + tid];
            // mode 2: every row reduces to its lower row and holds finite values
            //         (with a finite volume and no positivity: the fast update)
            // mode 1: every row reduces to its lower row
            // mode 0: some row needs the 4-tap interpolation
            int sh = 1, fin = 1;
            for (int r = 0; r < nr; ++r) {
                const TjBpRow row = rows[(size_t)(s0 + tid) * g.vh + j0 + r];
                if (row.valid && !row.shortcut) sh = 0;
                if (row.valid && !row.finite) fin = 0;
            }
            s_mode[tid] = sh ? ((fin && finite_volume && !POS) ? 2 : 1) : 0;
        }
        if (tid < G * R) {
            const int gg = tid / R, r = tid % R;
            int fl = 0, yb = 0, y1 = 0;
            double yf = 0.0;
            if (gg < gn && r < nr) {
                const TjBpRow row = rows[(size_t)(s0 + gg) * g.vh + j0 + r];
                fl = (row.valid ? 1 : 0) | (row.shortcut ? 2 : 0);
                yb = row.yb;
                y1 = row.y1;
                yf = row.yf;
            }
            s_rflag[gg][r] = fl;
            s_ryb[gg][r] = yb;
            s_ry1[gg][r] = y1;
            s_ryf[gg][r] = yf;
        }
        __syncthreads();
        // wave w stages image rows w, w+4, ... of the group straight from the
        // pair image into LDS: each 64-cell copy is one range-checked 16-byte
        // buffer load per lane (cells past the row, rows that take no value
        // and the zero cell XW-1 read 0); the last copy of a row is shifted
        // back to end at XW, rewriting cells of the previous one with the
        // same values, so no copy writes past the row
        {
            const int lane = tid & 63;
            const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
            for (int gr = wv; gr < gn * R; gr += TJ_BLOCK / 64) {
                const int gg = gr / R, r = gr - (gr / R) * R;
                const int x0s = __builtin_amdgcn_readfirstlane(s_x0[gg]);
                const bool rok = (__builtin_amdgcn_readfirstlane(s_rflag[gg][r]) & 1) != 0;
                const int ryb = __builtin_amdgcn_readfirstlane(s_ryb[gg][r]);
                const double2* prow = pairs + ((size_t)(s0 + gg) * g.ph + (rok ? ryb : 0)) * g.pw;
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void*)prow, 0, rok ? g.pw * 16 : 0, 0x00020000);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int c0 = min(64 * b, XW - 64);
                    const int xx = c0 + lane;
                    const uint32_t voff = (xx < XW - 1) ? (uint32_t)(x0s + xx) * 16u : 0xFFFFFFF0u;
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)&lds[gg][r][c0],
                                                             16, voff, 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);  // this thread's copies have landed
        __syncthreads();
        for (int gg = 0; gg < gn; ++gg) {
            const double e2 = s_e2[gg];
            const double xe = x * s_e0[gg];
            const int x0s = s_x0[gg];
            const int mode = s_mode[gg];
            if (mode == 2) {
                // finite images, finite volume, no positivity: lanes off the image
                // (and invalid rows) read zero pairs, and acc + 0 == acc
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    const double z = (k0 + kk) - g.vcz;
                    const double xx = (xe + z * e2) + g.pc2;
                    const bool vx = (xx >= 0) && (xx < pwd);
                    const double xs = vx ? xx : 0.0;
                    const int xb = (int)xs;
                    const double xf = xs - (double)xb;
                    const int li = vx ? xb - x0s : XW - 1;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const double2 q = lds[gg][r][li];
                        float a = acc[kk][r] + (float)(q.x + xf * q.y);
                        acc[kk][r] = (isnan(a) || isinf(a)) ? 0.0f : a;
                    }
                }
            } else if (mode == 1) {
                bool rv[R];
#pragma unroll
                for (int r = 0; r < R; ++r) rv[r] = (s_rflag[gg][r] & 1) != 0;
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    const double z = (k0 + kk) - g.vcz;
                    const double xx = (xe + z * e2) + g.pc2;
                    const bool vx = active && (kk < nk) && (xx >= 0) && (xx < pwd);
                    const double xs = vx ? xx : 0.0;
                    const int xb = (int)xs;
                    const double xf = xs - (double)xb;
                    const int li = vx ? xb - x0s : XW - 1;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const double2 q = lds[gg][r][li];
                        const double lower = q.x + xf * q.y;
                        // lower + 0*(upper - lower): no update unless lower is finite
                        float a = acc[kk][r] + (float)lower;
                        a = (isnan(a) || isinf(a)) ? 0.0f : a;
                        if (POS) a = (a < 0) ? 0.0f : a;
                        acc[kk][r] = (vx && rv[r] && isfinite(lower)) ? a : acc[kk][r];
                    }
                }
            } else {
                const float* img = images + (size_t)s_img[gg] * npx;
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    const double z = (k0 + kk) - g.vcz;
                    const double xx = (xe + z * e2) + g.pc2;
                    const bool vx = active && (kk < nk) && (xx >= 0) && (xx < pwd);
                    if (!vx) continue;
                    const int xb = (int)xx;
                    const double xf = xx - (double)xb;
                    const int li = xb - x0s;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if (!(s_rflag[gg][r] & 1)) continue;
                        const double2 q = lds[gg][r][li];
                        const double lower = q.x + xf * q.y;
                        float value;
                        if (s_rflag[gg][r] & 2) {
                            value = isfinite(lower) ? (float)lower : __builtin_nanf("");
                        } else {
                            const float* up = img + (size_t)s_ry1[gg][r] * g.pw;
                            const double ul = up[xb], ur = up[min(xb + 1, g.pw - 1)];
                            const double upper = ul + xf * (ur - ul);
                            value = (float)(lower + s_ryf[gg][r] * (upper - lower));
                        }
                        bp_accumulate<POS>(acc[kk][r], value);
                    }
                }
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (kk < nk && r < nr) vol[(size_t)(j0 + r) * plane + (size_t)(k0 + kk) * g.vw + i] = acc[kk][r];
}

// Fast backprojection for the common case of a launch: a finite volume, no
// positivity, and every image row a voxel row meets finite with an integer
// position (k_bp_rows leaves *slow clear).  Same results as k_bp_lds's finite
// mode with fewer instructions per update and the staging hidden:
// - the staged cells start at floor(min xx) over the block, which may be
//   negative, and cover every column the block's voxels reach; cells of
//   columns outside the image (and whole rows outside the detector) are
//   loaded as zero pairs by the buffer range check, so a voxel reads the cell
//   of floor(xx) with no range test: off the image it adds (0 + f*0) = 0,
//   exactly TomoJ's "skip" for a finite sum (a voxel holding -0 becomes +0);
// - for xx >= 0, floor(xx) is (int)xx and xx - floor(xx) is exact, so the
//   weight is ImageJ's x - (int)x;
// - one image at a time, double-buffered: the rows of image s+1 stream into
//   LDS (buffer-to-LDS loads) while image s is added;
// - slices past the volume in the last chunk read the last slice's cells and
//   are never stored.
// The per-image NaN/Inf scrub of TomoJ stays (a finite sum can overflow).
template <int R, int KC, int WAVES>
__global__ __launch_bounds__(TJ_BLOCK, WAVES) void k_bp_fast(
    TjGeomDev g, float* __restrict__ vol, const TjProj* __restrict__ pp, const int* __restrict__ slot_proj,
    int nslot, const TjBpRow* __restrict__ rows, int start_y, int end_y, const double2* __restrict__ pairs,
    const int* __restrict__ slow)
{
    if (*slow) return;  // the general kernel takes this launch (uniform)
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((int*)slow + 1, 1);  // launches taken (slow[1])
    constexpr int XW = TJ_BLOCK + KC + 8;
    constexpr int NB = (XW + 63) / 64;
    constexpr int NW = TJ_BLOCK / 64;
    static_assert(R <= NW, "one staging wave per row");
    __shared__ double2 lds[2][R][XW];
    __shared__ double s_e0[2], s_e2[2];
    __shared__ int s_x0[2];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nbx = (g.vw + TJ_BLOCK - 1) / TJ_BLOCK, nby = (end_y - start_y + R - 1) / R;
    const int nbz = (g.vd + KC - 1) / KC;
    int bx, by, bz;
    {
        // XCD-aware order as in k_bp_lds: a row group's blocks on one XCD
        const int b = blockIdx.x;
        if ((nby & 7) == 0) {
            const int per = nbx * nbz, k = b >> 3;
            by = (b & 7) + 8 * (k / per);
            const int rr = k - (k / per) * per;
            bz = rr / nbx;
            bx = rr - bz * nbx;
        } else {
            bx = b % nbx;
            by = (b / nbx) % nby;
            bz = b / (nbx * nby);
        }
    }
    const int i0 = bx * TJ_BLOCK;
    const int i = i0 + tid;
    const int j0 = start_y + by * R;
    const int nr = min(R, end_y - j0);
    const int k0 = bz * KC;
    const int nk = min(KC, g.vd - k0);
    const int ilast = min(i0 + TJ_BLOCK - 1, g.vw - 1);
    const bool active = i < g.vw;
    const double x = (active ? i : ilast) - g.vcx;
    const size_t plane = (size_t)g.vd * g.vw;

    // wave r (< nr) stages image row r of slot s into buffer b: the cells
    // from floor(min xx) on, one range-checked 16-byte buffer load per lane
    // and 64-cell copy; lane 0 of wave 0 leaves the slot's constants for the
    // compute phase
    auto stage = [&](int sl, int b) {
        const TjProj& Q = pp[slot_proj[sl]];
        const double xa = (i0 - g.vcx) * Q.e[0], xb = (ilast - g.vcx) * Q.e[0];
        const double za = (k0 - g.vcz) * Q.e[2], zb = (k0 + nk - 1 - g.vcz) * Q.e[2];
        const double m = fmin(fmin((xa + za) + g.pc2, (xa + zb) + g.pc2), fmin((xb + za) + g.pc2, (xb + zb) + g.pc2));
        const int x0s = __builtin_amdgcn_readfirstlane((int)floor(m));
        if (wv == 0 && lane == 0) {
            s_x0[b] = x0s;
            s_e0[b] = Q.e[0];
            s_e2[b] = Q.e[2];
        }
        if (wv < R) {
            const int r = wv;
            bool rok = false;
            int ryb = 0;
            if (r < nr) {
                const TjBpRow row = rows[(size_t)sl * g.vh + j0 + r];
                rok = row.valid != 0;
                ryb = row.yb;
            }
            rok = __builtin_amdgcn_readfirstlane(rok ? 1 : 0) != 0;
            ryb = __builtin_amdgcn_readfirstlane(ryb);
            const double2* prow = pairs + ((size_t)sl * g.ph + (rok ? ryb : 0)) * g.pw;
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc((void*)prow, 0, rok ? g.pw * 16 : 0, 0x00020000);
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int c0 = min(64 * c, XW - 64);
                const int col = x0s + c0 + lane;
                const uint32_t voff = (col >= 0) ? (uint32_t)col * 16u : 0xFFFFFFF0u;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)&lds[b][r][c0],
                                                         16, voff, 0, 0, 0);
            }
        }
    };

    if (nslot > 0) stage(0, 0);
    float acc[KC][R];
#pragma unroll
    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[kk][r] = (active && kk < nk && r < nr)
                             ? vol[(size_t)(j0 + r) * plane + (size_t)(k0 + kk) * g.vw + i]
                             : 0.0f;
    double zk[KC];  // slice positions, the last real slice for the chunk's padding
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) zk[kk] = (k0 + min(kk, nk - 1)) - g.vcz;

    for (int sl = 0; sl < nslot; ++sl) {
        const int b = sl & 1;
        __builtin_amdgcn_s_waitcnt(0);  // this thread's copies of image sl have landed
        __syncthreads();                // every thread's have, and image sl-1 is done
        if (sl + 1 < nslot) stage(sl + 1, b ^ 1);
        const double e2 = s_e2[b];
        const double xe = x * s_e0[b];
        const int x0s = s_x0[b];
        int c[KC];
        double xf[KC];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            const double xx = (xe + zk[kk] * e2) + g.pc2;
            const double fl = floor(xx);
            xf[kk] = xx - fl;
            c[kk] = (int)fl - x0s;
        }
        double2 q[2][R];
#pragma unroll
        for (int r = 0; r < R; ++r) q[0][r] = lds[b][r][c[0]];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            if (kk + 1 < KC) {
#pragma unroll
                for (int r = 0; r < R; ++r) q[(kk + 1) & 1][r] = lds[b][r][c[kk + 1 < KC ? kk + 1 : kk]];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double2 v = q[kk & 1][r];
                const float a = acc[kk][r] + (float)(v.x + xf[kk] * v.y);
                acc[kk][r] = (isnan(a) || isinf(a)) ? 0.0f : a;
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (kk < nk && r < nr) vol[(size_t)(j0 + r) * plane + (size_t)(k0 + kk) * g.vw + i] = acc[kk][r];
}

// Norm of every ray of a single-axis projection: with e01 = 0 the x(z) sequence
// of a ray does not depend on its detector row, so norm_sum
// (VoxelProjector3D.java:774-784) is a function of (projection, column) only.
// One thread per (projection, column), same loop as projectJosephPartial.
__global__ __launch_bounds__(TJ_BLOCK) void k_norm_table(TjGeomDev g, const TjProj* __restrict__ pp,
                                                         float* __restrict__ tab)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= g.pw) return;
    const TjProj& P = pp[p];
    float ns = 0.0f;
    if (P.axis) {
        const double ic = (i - g.pc0) / g.scale;
        double xc = ic - P.yfx * 0.0 - P.zfx * (-1 - g.vcz);
        xc /= P.divx;
        bool entered = false;
        for (int z = 0; z < g.vd; ++z) {
            xc = xc + P.dirx;
            const double x = xc + g.vcx;
            if (x + TJ_ACCURACY >= 0 && x < g.bxmax) {
                ns = (float)((double)ns + P.norminc);
                entered = true;
            } else if (entered) {
                ns = (float)((double)ns + P.norminc);
                break;
            }
        }
    }
    tab[(size_t)p * g.pw + i] = ns;
}

// Forward projection for a single vertical tilt axis with integer volume rows:
// a 256-thread block owns 4 consecutive detector rows (4 consecutive volume
// rows), 256 detector columns and G added projections (thread = column, one ray
// per projection).  The 4 volume rows are staged in LDS interleaved, one float4
// per (z, x), with column W a copy of column W-1, so the two taps of all four
// rows come from two 16-byte reads and the nearest-column rule of getPixel
// needs no branch (lerp with equal taps).  Lanes outside the volume read a
// zero cell.  Requires a finite volume and rows whose y is an integer (the
// host checks both; otherwise k_fp_axis runs).  Blocks are numbered so the
// blocks sharing one row group run on one XCD back to back and share its L2.
// NG: the rays of the thread that belong to projections of the launch (the
// last group of a launch can hold fewer than G); only they are marched.
//
// The float4 at LDS byte address a.
__device__ __forceinline__ float4 lds_f4(uint32_t a)
{
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v t = *(const TJ_LDS f4v*)a;
    return make_float4(t.x, t.y, t.z, t.w);
}

// One slice of rays: the same arithmetic for both staging forms below.
template <int G, int NG, bool NEARY, bool LAST>
__device__ __forceinline__ void fp_lds_slice(const TjGeomDev& g, const float4* lds4, int zo, int zero_idx,
                                             const float (&ym)[4], double (&xc)[G], const double (&dirx)[G],
                                             float4 (&rs)[G])
{
    const double nacc = -TJ_ACCURACY;  // x + ACCURACY >= 0  <=>  x >= -ACCURACY
    const double bx = g.bxmax;
    const double wm1 = g.vw - 1;
    // LDS byte address of the slice's first cell, kept in a scalar register, so
    // a cell address is one shift-add of its index; the zero cell counted from
    // the slice
    uint32_t sz = __builtin_amdgcn_readfirstlane((uint32_t)(size_t)(const TJ_LDS float4*)(lds4 + zo));
    asm("" : "+s"(sz));
    const int zrel = zero_idx - zo;
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        xc[k] = xc[k] + dirx[k];
        const double xv = xc[k] + g.vcx;
        const bool in = (xv >= nacc) && (xv < bx);
        // inside, xv is in [-ACCURACY, W): clamping to [0, W-1] keeps xb and,
        // for xb < W-1, the fraction; at xb = W-1 both taps are equal (the
        // replicated column) so the fraction does not matter
        const double xs = fmin(fmax(xv, 0.0), wm1);
        const int xb = (int)xs;
        // the last slice takes the nearest value (getPixel's z == D-1 rule):
        // its fraction times 0
        float xf = (float)xs - (float)xb;
        if (LAST) xf = xf * 0.0f;
        const uint32_t a = sz + 16u * (uint32_t)(in ? xb : zrel);
        const float4 q0 = lds_f4(a), q1 = lds_f4(a + 16u);
        if (NEARY) {
            rs[k].x += q0.x + (xf * ym[0]) * (q1.x - q0.x);
            rs[k].y += q0.y + (xf * ym[1]) * (q1.y - q0.y);
            rs[k].z += q0.z + (xf * ym[2]) * (q1.z - q0.z);
            rs[k].w += q0.w + (xf * ym[3]) * (q1.w - q0.w);
        } else {
            rs[k].x += q0.x + xf * (q1.x - q0.x);
            rs[k].y += q0.y + xf * (q1.y - q0.y);
            rs[k].z += q0.z + xf * (q1.z - q0.z);
            rs[k].w += q0.w + xf * (q1.w - q0.w);
        }
    }
}

// Double-buffered staging (PF): the slab alternates between two LDS buffers of
// zc * (W + 1) + 2 cells (the slab and its two zero cells).  The loads of
// chunk c + 1 are issued into registers before chunk c is projected and
// stored into the other buffer after it, so their latency hides behind the
// rays and a chunk needs one barrier instead of two.  Thread t owns the cells
// t + q * (threads of the block), q < TJ_PFQ, of every chunk: the host picks zc
// with zc * (W + 1) <= TJ_PFQ * threads.  The cells and the arithmetic are the
// single-buffer form's, so the sums are identical.
#define TJ_PFQ 4
template <int G, int NG, bool NEARY, int GB>
__device__ __forceinline__ void fp_lds_chunks_pf(const TjGeomDev& g, const float* __restrict__ vol, float4* lds4,
                                                 int zc, int rowbase, const float (&ym)[4], double (&xc)[G],
                                                 const double (&dirx)[G], float4 (&rs)[G])
{
    constexpr int NT = TJ_BLOCK * GB;
    const int tid = threadIdx.x;
    const size_t plane = (size_t)g.vd * g.vw;
    const int W1 = g.vw + 1;
    const int bufsz = zc * W1 + 2;
    const int zero_idx = zc * W1;
    const bool rin0 = rowbase + 0 >= 0 && rowbase + 0 < g.vh, rin1 = rowbase + 1 >= 0 && rowbase + 1 < g.vh;
    const bool rin2 = rowbase + 2 >= 0 && rowbase + 2 < g.vh, rin3 = rowbase + 3 >= 0 && rowbase + 3 < g.vh;
    const float* r0 = vol + (size_t)(rin0 ? rowbase + 0 : 0) * plane;
    const float* r1 = vol + (size_t)(rin1 ? rowbase + 1 : 0) * plane;
    const float* r2 = vol + (size_t)(rin2 ? rowbase + 2 : 0) * plane;
    const float* r3 = vol + (size_t)(rin3 ? rowbase + 3 : 0) * plane;
    // the thread's cells: slice zq within the chunk, offset oq within the
    // chunk's planes (columns past W read column W-1); zq = zc for none
    int zq[TJ_PFQ], oq[TJ_PFQ];
#pragma unroll
    for (int q = 0; q < TJ_PFQ; ++q) {
        const int e = tid + q * NT;
        const int z = e / W1;
        zq[q] = min(z, zc);
        oq[q] = z * g.vw + min(e - z * W1, g.vw - 1);
    }
    float v[TJ_PFQ][4];
    auto load = [&](int z0) {
        const int zn = min(zc, g.vd - z0);
        const int zo = z0 * g.vw;
#pragma unroll
        for (int q = 0; q < TJ_PFQ; ++q)
            if (zq[q] < zn) {
                const int o = zo + oq[q];
                v[q][0] = r0[o];
                v[q][1] = r1[o];
                v[q][2] = r2[o];
                v[q][3] = r3[o];
            }
    };
    auto store = [&](float4* dst, int z0) {
        const int zn = min(zc, g.vd - z0);
#pragma unroll
        for (int q = 0; q < TJ_PFQ; ++q)
            if (zq[q] < zn)
                dst[tid + q * NT] = make_float4(rin0 ? v[q][0] : 0.0f, rin1 ? v[q][1] : 0.0f,
                                                rin2 ? v[q][2] : 0.0f, rin3 ? v[q][3] : 0.0f);
    };
    load(0);
    store(lds4, 0);
    __syncthreads();
    int buf = 0;
    for (int z0 = 0; z0 < g.vd; z0 += zc) {
        const int zn = min(zc, g.vd - z0);
        const bool more = z0 + zc < g.vd;  // uniform
        if (more) load(z0 + zc);
        const float4* cur = lds4 + buf * bufsz;
        // the last slice takes the nearest value (getPixel's z == D-1 rule)
        const int zlast = more ? zn : zn - 1;
        for (int zz = 0; zz < zlast; ++zz)
            fp_lds_slice<G, NG, NEARY, false>(g, cur, zz * W1, zero_idx, ym, xc, dirx, rs);
        if (zlast < zn) fp_lds_slice<G, NG, NEARY, true>(g, cur, zlast * W1, zero_idx, ym, xc, dirx, rs);
        if (more) {
            store(lds4 + (buf ^ 1) * bufsz, z0 + zc);
            __syncthreads();
            buf ^= 1;
        }
    }
}

template <int G, int NG, bool NEARY, int GB>
__device__ __forceinline__ void fp_lds_chunks(const TjGeomDev& g, const float* __restrict__ vol, float4* lds4,
                                              int zc, int rowbase, int ns, const float (&ym)[4], double (&xc)[G],
                                              const double (&dirx)[G], float4 (&rs)[G])
{
    const int tid = threadIdx.x;
    const size_t plane = (size_t)g.vd * g.vw;
    const int W1 = g.vw + 1;
    const int zero_idx = zc * W1;
    const bool rin0 = rowbase + 0 >= 0 && rowbase + 0 < g.vh, rin1 = rowbase + 1 >= 0 && rowbase + 1 < g.vh;
    const bool rin2 = rowbase + 2 >= 0 && rowbase + 2 < g.vh, rin3 = rowbase + 3 >= 0 && rowbase + 3 < g.vh;
    const float* r0 = vol + (size_t)(rin0 ? rowbase + 0 : 0) * plane;
    const float* r1 = vol + (size_t)(rin1 ? rowbase + 1 : 0) * plane;
    const float* r2 = vol + (size_t)(rin2 ? rowbase + 2 : 0) * plane;
    const float* r3 = vol + (size_t)(rin3 ? rowbase + 3 : 0) * plane;
    const int lane = tid & 63, wave = tid >> 6;
    for (int z0 = 0; z0 < g.vd; z0 += zc) {
        const int zn = min(zc, g.vd - z0);
        __syncthreads();
        // wave w stages slices w, w + (waves of the block), ...; lanes take
        // consecutive columns (32-bit offsets within a plane: the plane D*W is
        // far below 2^31)
        // SU columns per lane per batch: 4*SU loads in flight before the first
        // LDS write.  Rows outside the volume read row 0 (a valid address) and
        // store 0; columns past W read column W-1.
        for (int zz = wave; zz < zn; zz += TJ_BLOCK * GB / 64) {
            const int zo = (z0 + zz) * g.vw;
            float4* dst = lds4 + zz * W1;
            for (int x0 = 0; x0 < W1; x0 += 64 * TJ_SU) {
                float v[TJ_SU][4];
#pragma unroll
                for (int u = 0; u < TJ_SU; ++u) {
                    const int o = zo + min(x0 + 64 * u + lane, g.vw - 1);
                    v[u][0] = r0[o];
                    v[u][1] = r1[o];
                    v[u][2] = r2[o];
                    v[u][3] = r3[o];
                }
#pragma unroll
                for (int u = 0; u < TJ_SU; ++u) {
                    const int xx = x0 + 64 * u + lane;
                    if (xx < W1)
                        dst[xx] = make_float4(rin0 ? v[u][0] : 0.0f, rin1 ? v[u][1] : 0.0f, rin2 ? v[u][2] : 0.0f,
                                              rin3 ? v[u][3] : 0.0f);
                }
            }
        }
        __syncthreads();
        // the last slice takes the nearest value (getPixel's z == D-1 rule): its
        // fraction is multiplied by 0
        const int zlast = (z0 + zn == g.vd) ? zn - 1 : zn;
        for (int zz = 0; zz < zlast; ++zz)
            fp_lds_slice<G, NG, NEARY, false>(g, lds4, zz * W1, zero_idx, ym, xc, dirx, rs);
        if (zlast < zn) fp_lds_slice<G, NG, NEARY, true>(g, lds4, zlast * W1, zero_idx, ym, xc, dirx, rs);
    }
}

// With FUSE the kernel also does k_residual's work for its pixels: instead of
// proj and norm it writes diff = (float)(((exp - th) / norm) * factor), the
// FP64 sum of the squared differences per (slot, 4-row block, column tile) and
// a non-finite flag per (slot, row, column tile); k_err_fold folds both.
struct TjFpFuse {
    const float* exps;  // [n][ph][pw], image slot_proj[s] for slot s
    double factor;
    float* diff;        // [nslot][ph][pw]
    double* rerr;       // [nslot][ceil(ph / 4)][ntile]
    uint8_t* rbad;      // [nslot][ph][ntile]
    float* proj;        // [nslot][ph][pw] or null: the projections and ray norms too,
    float* norm;        //   as the unfused kernel writes them (getProjection)
};

// GB: groups of G projections per block, one 256-thread part each; the parts
// share the staged slab (the slab is read from L2 once per GB G projections).
// The last set of a launch may hold parts without projections: they stage and
// wait at the barriers like the others and march no ray.
template <int G, bool FUSE, int GB, bool PF = false>
__global__ __launch_bounds__(TJ_BLOCK * GB) void k_fp_lds(
    TjGeomDev g, const float* __restrict__ vol, const TjProj* __restrict__ pp,
    const int* __restrict__ slot_proj, int nslot, int start_row, const float* __restrict__ norm_tab, int zc,
    int nrb, int nsub, int ntile, float* __restrict__ proj, float* __restrict__ norm, TjFpFuse fu)
{
    extern __shared__ float4 lds4[];
    const int tid = threadIdx.x & (TJ_BLOCK - 1);   // column of the tile
    const int part = GB > 1 ? (int)threadIdx.x / TJ_BLOCK : 0;
    const int b = blockIdx.x;
    const int xcd = b & 7, local = b >> 3;
    const int sub = local % nsub;
    const int rb = (local / nsub) * 8 + xcd;
    if (rb >= nrb) return;  // whole block
    const int tile = sub % ntile, grp = sub / ntile;
    const int i = tile * TJ_BLOCK + tid;
    const int s0 = (grp * GB + part) * G;
    const int ns = min(G, nslot - s0);  // <= 0: a part without projections
    const int s0c = min(s0, nslot - 1);
    const int j0 = rb * 4;
    const bool col = i < g.pw;
    const double ic = ((col ? i : g.pw - 1) - g.pc0) / g.scale;

    // the rows a block stages do not depend on the projection (the host admits
    // only launches whose projections share the row mapping)
    const TjProj& P0 = pp[slot_proj[s0c]];
    bool ok[4];
    float ym[4];  // 0 for the last volume row (nearest value), 1 otherwise
    int rowbase = 0;
    bool have_row = false;
    double oldy_any = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        ok[r] = false;
        ym[r] = 1.0f;
        const int j = j0 + r;
        if (j < g.ph && j >= start_row) {
            const double jc = jcentered(g, start_row, j);
            const double oldy = ((jc - P0.icf * ic) - P0.zcf * (-1 - g.vcz)) / P0.divy;
            oldy_any = oldy;
            const double y = (oldy + P0.diry) + g.vcy;
            if (y + TJ_ACCURACY >= 0 && y < g.bymax) {
                const int yb = (y < 0) ? 0 : (int)y;
                ok[r] = true;
                ym[r] = (yb == g.vh - 1) ? 0.0f : 1.0f;
                if (!have_row) {
                    rowbase = yb - r;
                    have_row = true;
                }
            }
        }
    }
    const bool neary = (ym[0] * ym[1] * ym[2] * ym[3]) == 0.0f;

    double xc[G], dirx[G];
    float4 rs[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        rs[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        const TjProj& Q = pp[slot_proj[min(s0 + min(k, max(ns, 1) - 1), nslot - 1)]];
        double ox = ic - Q.yfx * oldy_any - Q.zfx * (-1 - g.vcz);
        ox /= Q.divx;
        // rays beyond the launch's projections stay far outside the volume
        xc[k] = (k < ns) ? ox : -1.0e300;
        dirx[k] = (k < ns) ? Q.dirx : 0.0;
    }
    if (have_row) {
        // the zero cells after the slab (of both buffers with PF)
        if (PF) {
            if (threadIdx.x < 4)
                lds4[(threadIdx.x >> 1) * (zc * (g.vw + 1) + 2) + zc * (g.vw + 1) + (threadIdx.x & 1)] =
                    make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (threadIdx.x < 2) {
            lds4[zc * (g.vw + 1) + threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        auto run = [&](auto ng) {
            constexpr int NG = decltype(ng)::value;
            if constexpr (PF) {
                if (neary)
                    fp_lds_chunks_pf<G, NG, true, GB>(g, vol, lds4, zc, rowbase, ym, xc, dirx, rs);
                else
                    fp_lds_chunks_pf<G, NG, false, GB>(g, vol, lds4, zc, rowbase, ym, xc, dirx, rs);
            } else {
                if (neary)
                    fp_lds_chunks<G, NG, true, GB>(g, vol, lds4, zc, rowbase, ns, ym, xc, dirx, rs);
                else
                    fp_lds_chunks<G, NG, false, GB>(g, vol, lds4, zc, rowbase, ns, ym, xc, dirx, rs);
            }
        };
        // the smallest supported count of marched rays that covers ns
        if (ns >= G) run(std::integral_constant<int, G>{});
        else if (ns == 1) run(std::integral_constant<int, 1>{});
        else if (ns == 2) run(std::integral_constant<int, (G > 2 ? 2 : G)>{});
        else if (ns == 3) run(std::integral_constant<int, (G > 3 ? 3 : G)>{});
        else if (ns == 4) run(std::integral_constant<int, (G > 4 ? 4 : G)>{});
        else run(std::integral_constant<int, G>{});
    }
    if constexpr (FUSE) {
        __shared__ double red[TJ_BLOCK * GB / 64][G];
        __shared__ unsigned redb[TJ_BLOCK * GB / 64][G * 4];
        const int lane = tid & 63, wv = (int)threadIdx.x >> 6;  // wave of the block
        const int w0 = part * (TJ_BLOCK / 64);                   // first wave of the part
        const size_t npx = (size_t)g.pw * g.ph;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k >= ns) break;  // uniform over the block
            const int p = slot_proj[s0 + k];
            const double ninc = pp[p].norminc;
            const float nt = col ? norm_tab[(size_t)p * g.pw + i] : 0.0f;
            const float rv[4] = {rs[k].x, rs[k].y, rs[k].z, rs[k].w};
            double v2 = 0.0;  // this lane's squared differences over the block's rows
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + r;
                if (j >= g.ph) break;  // uniform over the block
                bool bad = false;
                if (col) {
                    float th, n;
                    if (j < start_row) {
                        th = 0.0f;
                        n = 0.0f;
                    } else if (ok[r] && nt > 0) {
                        th = (float)((double)rv[r] * ninc);
                        n = nt;
                    } else {
                        th = __builtin_nanf("");
                        n = 0.0f;
                    }
                    const float ex = fu.exps[(size_t)p * npx + (size_t)j * g.pw + i];
                    float d = 0.0f;
                    if (!isnan(ex) && !isnan(th)) {
                        const double value = (double)((ex - th) / n) * fu.factor;
                        d = (float)value;
                        v2 += value * value;
                    }
                    const size_t o = ((size_t)(s0 + k) * g.ph + j) * g.pw + i;
                    fu.diff[o] = d;
                    if (fu.proj) {  // uniform
                        fu.proj[o] = th;
                        fu.norm[o] = n;
                    }
                    bad = !isfinite(d);
                }
                const unsigned wbad = __ballot(bad) != 0 ? 1u : 0u;
                if (lane == 0) redb[wv][k * 4 + r] = wbad;
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) v2 += __shfl_xor(v2, m);
            if (lane == 0) red[wv][k] = v2;
        }
        __syncthreads();
        if (tid < ns * 4) {
            const int k = tid >> 2, r = tid & 3, j = j0 + r;
            if (j < g.ph) {
                unsigned b = 0;
#pragma unroll
                for (int w = 0; w < TJ_BLOCK / 64; ++w) b |= redb[w0 + w][tid];
                fu.rbad[((size_t)(s0 + k) * g.ph + j) * ntile + tile] = (uint8_t)b;
            }
            if (r == 0) {
                double sum = 0.0;
#pragma unroll
                for (int w = 0; w < TJ_BLOCK / 64; ++w) sum += red[w0 + w][k];
                fu.rerr[((size_t)(s0 + k) * nrb + rb) * ntile + tile] = sum;
            }
        }
        return;
    }
    if (!col) return;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k >= ns) break;
        const int p = slot_proj[s0 + k];
        const double ninc = pp[p].norminc;
        const float nt = norm_tab[(size_t)p * g.pw + i];
        const float rv[4] = {rs[k].x, rs[k].y, rs[k].z, rs[k].w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = j0 + r;
            if (j >= g.ph) break;
            const size_t o = ((size_t)(s0 + k) * g.ph + j) * g.pw + i;
            if (j < start_row) {
                proj[o] = 0.0f;
                norm[o] = 0.0f;
            } else if (ok[r] && nt > 0) {
                proj[o] = (float)((double)rv[r] * ninc);
                norm[o] = nt;
            } else {
                proj[o] = __builtin_nanf("");
                norm[o] = 0.0f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Projection supply with the identity alignment (TiltSeries.getPixels ->
// AffineAlignment.applyTransformOnImage, AffineAlignment.java:136-156): each
// image is normalised by the mean and sample standard deviation of its
// tilt-dependent ROI (TiltSeries.getImageStatistics :809-835; columns
// [x0, x1), all rows; ImageJ FloatStatistics: double sums, stdDev =
// sqrt(((n*sum2 - sum^2)/n)/(n-1))).  One block per image, fixed reduction tree.
// ---------------------------------------------------------------------------
// roi: 4 ints per image, the pixel rectangle [x0, x1) x [y0, y1) (clipped)
__global__ __launch_bounds__(TJ_BLOCK) void k_roi_stats(const float* __restrict__ imgs, int pw, int ph,
                                                        const int* __restrict__ roi, double* __restrict__ stats)
{
    __shared__ double sh[TJ_BLOCK];
    const int p = blockIdx.x;
    const int x0 = max(roi[4 * p], 0), x1 = min(roi[4 * p + 1], pw);
    const int y0 = max(roi[4 * p + 2], 0), y1 = min(roi[4 * p + 3], ph);
    const int rw = max(x1 - x0, 0);
    const size_t cnt = (size_t)rw * max(y1 - y0, 0);
    const float* img = imgs + (size_t)p * pw * ph;
    double sum = 0.0, sum2 = 0.0;
    for (size_t k = threadIdx.x; k < cnt; k += TJ_BLOCK) {
        const int y = y0 + (int)(k / rw), x = x0 + (int)(k - (size_t)(y - y0) * rw);
        const double v = img[(size_t)y * pw + x];
        sum += v;
        sum2 += v * v;
    }
    sum = block_reduce(sum, sh, AddD());
    sum2 = block_reduce(sum2, sh, AddD());
    if (threadIdx.x == 0) {
        const double n = (double)cnt;
        double sd = 0.0;
        if (n > 0) {
            sd = (n * sum2 - sum * sum) / n;
            sd = (sd > 0.0) ? sqrt(sd / (n - 1.0)) : 0.0;
        }
        stats[2 * p] = n > 0 ? sum / n : 0.0;
        stats[2 * p + 1] = sd;
    }
}

// Statistics over getImageStatistics' polygon ROI when the alignment rotates
// or shears it (TiltSeries.java:809-835): poly = 8 ints per image (xs[4],
// ys[4]).  A pixel counts when its centre lies inside by the even-odd crossing
// test (ImageJ's documented polygon-mask rule); bounding box clipped to the
// image; mean and sample deviation as k_roi_stats.
__device__ inline bool poly_inside(const int* xs, const int* ys, double px, double py)
{
    bool in = false;
    for (int i = 0, j = 3; i < 4; j = i++) {
        const double xi = xs[i], yi = ys[i], xj = xs[j], yj = ys[j];
        if ((yi > py) != (yj > py) && px < (xj - xi) * (py - yi) / (yj - yi) + xi) in = !in;
    }
    return in;
}

__global__ __launch_bounds__(TJ_BLOCK) void k_poly_stats(const float* __restrict__ imgs, int pw, int ph,
                                                         const int* __restrict__ poly, double* __restrict__ stats)
{
    __shared__ double sh[TJ_BLOCK];
    const int p = blockIdx.x;
    int xs[4], ys[4];
    for (int k = 0; k < 4; ++k) {
        xs[k] = poly[8 * p + k];
        ys[k] = poly[8 * p + 4 + k];
    }
    const int x0 = max(min(min(xs[0], xs[1]), min(xs[2], xs[3])), 0);
    const int x1 = min(max(max(xs[0], xs[1]), max(xs[2], xs[3])), pw);
    const int y0 = max(min(min(ys[0], ys[1]), min(ys[2], ys[3])), 0);
    const int y1 = min(max(max(ys[0], ys[1]), max(ys[2], ys[3])), ph);
    const int rw = max(x1 - x0, 0);
    const size_t cnt = (size_t)rw * max(y1 - y0, 0);
    const float* img = imgs + (size_t)p * pw * ph;
    double sum = 0.0, sum2 = 0.0, n = 0.0;
    for (size_t k = threadIdx.x; k < cnt; k += TJ_BLOCK) {
        const int y = y0 + (int)(k / rw), x = x0 + (int)(k - (size_t)(y - y0) * rw);
        if (!poly_inside(xs, ys, x + 0.5, y + 0.5)) continue;
        const double v = img[(size_t)y * pw + x];
        sum += v;
        sum2 += v * v;
        n += 1.0;
    }
    sum = block_reduce(sum, sh, AddD());
    sum2 = block_reduce(sum2, sh, AddD());
    n = block_reduce(n, sh, AddD());
    if (threadIdx.x == 0) {
        double sd = 0.0;
        if (n > 0) {
            sd = (n * sum2 - sum * sum) / n;
            sd = (sd > 0.0) ? sqrt(sd / (n - 1.0)) : 0.0;
        }
        stats[2 * p] = n > 0 ? sum / n : 0.0;
        stats[2 * p + 1] = sd;
    }
}

__global__ __launch_bounds__(TJ_BLOCK) void k_normalize(float* __restrict__ imgs, size_t npx,
                                                        const double* __restrict__ stats)
{
    const int p = blockIdx.y;
    const double mean = stats[2 * p], sd = stats[2 * p + 1];
    float* img = imgs + (size_t)p * npx;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < npx; k += (size_t)gridDim.x * blockDim.x)
        img[k] = (float)((img[k] - mean) / sd);
}

// ---------------------------------------------------------------------------
// Aligned projection supply: AffineAlignment.applyTransformOnImage for a
// non-identity transform (AffineAlignment.java:157-254).  Pixel (i, j) of image
// p samples the original at Tinv (i - w/2, j - h/2) + (w/2, h/2), evaluated as
// java.awt.geom.AffineTransform.transform does (m00 x + m01 y + m02, in FP64,
// no fused multiply-add); (int) truncation toward zero, so positions in
// (-1, 0) sample column or row 0 with a negative fraction as Java does; the
// last column and row take the nearest value; the bilinear weights are
// Java's (1 - dx - dy + dx dy, dx - dx dy, dy - dx dy, dx dy) in FP64.
// Normalisation (float)((pix - mean) / stdDev) after the cast; pixels that
// map outside take the fill: mean (FILL_AVG, not normalised), 0 (FILL_AVG
// normalised, FILL_NONE) or NaN (FILL_NaN).  inv: 6 doubles per image in the
// order m00, m10, m01, m11, m02, m12.
// ---------------------------------------------------------------------------
// ident[p] != 0: the identity branch (AffineAlignment.java:148-156), a copy or
// (float)((x - mean) / stdDev) of the pixel itself.
__global__ __launch_bounds__(TJ_BLOCK) void k_affine(const float* __restrict__ src, float* __restrict__ dst,
                                                     int w, int h, const double* __restrict__ inv,
                                                     const int* __restrict__ ident,
                                                     const double* __restrict__ stats, int normalize, int fill)
{
    const int i = blockIdx.x * TJ_BLOCK + threadIdx.x;
    const int j = blockIdx.y, p = blockIdx.z;
    if (i >= w) return;
    if (ident[p]) {
        const size_t o = (size_t)p * w * h + (size_t)j * w + i;
        dst[o] = normalize ? (float)((src[o] - stats[2 * p]) / stats[2 * p + 1]) : src[o];
        return;
    }
    const double* m = inv + 6 * p;
    const double cx = w / 2.0, cy = h / 2.0;
    const double x = i - cx, y = j - cy;
    const double xx = (m[0] * x + m[2] * y + m[4]) + cx;
    const double yy = (m[1] * x + m[3] * y + m[5]) + cy;
    const int ix0 = (int)xx, iy0 = (int)yy;
    const double dx0 = xx - ix0, dy0 = yy - iy0;
    const float* img = src + (size_t)p * w * h;
    const double mean = stats ? stats[2 * p] : 0.0, sd = stats ? stats[2 * p + 1] : 1.0;
    float out;
    if (ix0 >= 0 && ix0 < w && iy0 >= 0 && iy0 < h) {
        float pix;
        if (ix0 == w - 1 || iy0 == h - 1) {
            pix = img[ix0 + (size_t)iy0 * w];
        } else {
            const double fac4 = dx0 * dy0;
            const size_t pos = ix0 + (size_t)iy0 * w;
            const double v1 = img[pos], v3 = img[pos + w], v2 = img[pos + 1], v4 = img[pos + 1 + w];
            pix = (float)(v1 * (1 - dx0 - dy0 + fac4) + v2 * (dx0 - fac4) + v3 * (dy0 - fac4) + v4 * fac4);
        }
        if (normalize) pix = (float)((pix - mean) / sd);
        out = pix;
    } else if (fill == 1) {
        out = normalize ? 0.0f : (float)mean;
    } else if (fill == 2) {
        out = NAN;
    } else {
        out = 0.0f;
    }
    dst[(size_t)p * w * h + (size_t)j * w + i] = out;
}

// ---------------------------------------------------------------------------
// FISTA extrapolation of FistaProjector3D.endOfIteration
// (FistaProjector3D.java:43-59) for every voxel: the float difference with the
// volume of the start of the iteration, scaled in double, added back as a
// float unless it is NaN (TomoReconstruction2.addToPixel :365-370).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TJ_BLOCK) void k_fista(float* __restrict__ vol, const float* __restrict__ prev,
                                                    size_t n, double factor)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const float x = vol[k];
        const float d = x - prev[k];
        const double val = factor * (double)d;
        if (!isnan(val)) vol[k] = x + (float)val;
    }
}

// ---------------------------------------------------------------------------
// Hysteresis threshold of the projection supply, TiltSeries.applyThreshold
// (TiltSeries.java:720-806).  k_thr_levels (one block per image): ImageJ
// FloatStatistics' min, max and 256-bin histogram over the values within
// +-FLT_MAX, the mean as Java's sequential double sum (one lane, only for the
// average fill), and the two levels.  k_thr_apply: below th1 -> fill value,
// [th1, th2) -> queued.  k_thr_grow: a queued pixel whose neighbour (index
// -+1, wrapping across rows as TomoJ's test does, and -+W) holds the fill value
// (NaN in NaN mode) takes it; repeated until nothing changes, which reaches the
// same closure as TomoJ's backward sweeps.
// ---------------------------------------------------------------------------
__device__ inline int java_d2i(double x)
{
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return -2147483647 - 1;
    return (int)x;
}

__device__ inline bool fsv_in(double v) { return v >= -3.4028234663852886e38 && v <= 3.4028234663852886e38; }

// lv[img] = {th1, th2, value bits, mean}
__global__ __launch_bounds__(TJ_BLOCK) void k_thr_levels(const float* __restrict__ imgs, int w, int h,
                                                         double thr, double hys, int fill, double* __restrict__ lv)
{
    __shared__ double shd[TJ_BLOCK];
    __shared__ int hist[256];
    const size_t len = (size_t)w * h;
    const float* im = imgs + (size_t)blockIdx.x * len;
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (size_t i = threadIdx.x; i < len; i += blockDim.x) {
        const double v = im[i];
        if (fsv_in(v)) {
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    }
    mn = block_reduce(mn, shd, MinD());
    mx = block_reduce(mx, shd, MaxD());
    for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    const double scale = 256 / (mx - mn);
    for (size_t i = threadIdx.x; i < len; i += blockDim.x) {
        const double v = im[i];
        if (fsv_in(v) && v >= mn && v <= mx) {
            int index = java_d2i(scale * (v - mn));
            if (index >= 256) index = 255;
            atomicAdd(&hist[index], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mean = 0.0;
        if (fill == 1) {
            double sum = 0.0;
            long cnt = 0;
            for (size_t i = 0; i < len; ++i) {
                const double v = im[i];
                if (fsv_in(v) && v >= mn && v <= mx) {
                    sum += v;
                    ++cnt;
                }
            }
            mean = sum / cnt;
        }
        const int pixnb1 = java_d2i((double)w * h * thr);
        const int pixnb2 = java_d2i((double)w * h * hys);
        double th1 = 0, th2 = 0;
        long run = 0;
        for (int i = 0; i < 256; ++i) {
            run += hist[i];
            if (run < pixnb1) th1 = i;
            if (run < pixnb2) th2 = i;
        }
        th1 /= 256;
        th1 = th1 * (mx - mn) + mn;
        th2 /= 256;
        th2 = th2 * (mx - mn) + mn;
        float value = __int_as_float(0x7fc00000);
        if (fill == 1) value = (float)mean;
        if (fill == 0) value = 0.0f;
        double* o = lv + 4 * (size_t)blockIdx.x;
        o[0] = th1;
        o[1] = th2;
        o[2] = (double)__float_as_uint(value);
        o[3] = mean;
    }
}

__global__ __launch_bounds__(TJ_BLOCK) void k_thr_apply(float* __restrict__ imgs, size_t len,
                                                        const double* __restrict__ lv, uint8_t* __restrict__ q)
{
    const size_t base = (size_t)blockIdx.y * len;
    const double* o = lv + 4 * (size_t)blockIdx.y;
    const double th1 = o[0], th2 = o[1];
    const float value = __uint_as_float((unsigned)o[2]);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        const double v = imgs[base + i];
        uint8_t qq = 0;
        if (v < th1) imgs[base + i] = value;
        else if (v < th2) qq = 1;
        q[base + i] = qq;
    }
}

__global__ __launch_bounds__(TJ_BLOCK) void k_thr_grow(float* __restrict__ imgs, size_t len, int w,
                                                       const double* __restrict__ lv, uint8_t* __restrict__ q,
                                                       int* __restrict__ changed)
{
    const size_t base = (size_t)blockIdx.y * len;
    const float value = __uint_as_float((unsigned)lv[4 * (size_t)blockIdx.y + 2]);
    const bool nan_mode = value != value;
    float* im = imgs + base;
    int hit_any = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        if (!q[base + i]) continue;
        auto match = [&](size_t j) {
            const float v = im[j];
            return nan_mode ? (v != v) : (v == value);
        };
        const bool hit = (i >= 1 && match(i - 1)) || (i + 1 < len && match(i + 1)) ||
                         (i >= (size_t)w && match(i - w)) || (i + w < len && match(i + w));
        if (hit) {
            im[i] = value;
            q[base + i] = 0;
            hit_any = 1;
        }
    }
    if (hit_any) *changed = 1;
}

// ---------------------------------------------------------------------------
// Total-variation denoising, TotalVariationDenoising_3D.denoise3D
// (Filters/.../TotalVariationDenoising_3D.java:84-202), on the device layout
// [y][z][x] with p held as three planes px, py, pz of the volume's size.  A
// block owns 256 columns x of one row y over a run of TJ_TV_ZR slices and
// marches z, carrying the next slice's u (updateP) or the previous slice's pz
// (updateU) in a register.  Blocks are ordered so that each XCD sweeps the rows
// y of its share in turn: the neighbouring row a block reads (y+1 in updateP,
// y-1 in updateU) is the row the XCD's previous blocks read, still in its L2.
// Float arithmetic in Java's order; this file is built without contraction and
// with correctly rounded division and square root.
// ---------------------------------------------------------------------------
#define TJ_TV_ZR 64

struct TvTile {
    int x, y, z0, z1;
};

__device__ inline bool tv_tile(int W, int H, int D, int nreal, TvTile& t)
{
    const int nb = gridDim.x;  // a multiple of 8
    const int b = blockIdx.x;
    const int k = (b & 7) * (nb >> 3) + (b >> 3);
    if (k >= nreal) return false;
    const int nxt = (W + TJ_BLOCK - 1) / TJ_BLOCK;
    const int r = k / nxt;
    t.x = (k - r * nxt) * TJ_BLOCK + (int)threadIdx.x;
    t.y = r % H;
    t.z0 = (r / H) * TJ_TV_ZR;
    t.z1 = min(t.z0 + TJ_TV_ZR, D);
    return t.x < W;
}

// updateP (:115-156): a = dt/theta, b = dt/theta/g, both rounded to float as
// Java forms them.
__global__ __launch_bounds__(TJ_BLOCK) void k_tv_p(const float* __restrict__ u, float* __restrict__ px,
                                                   float* __restrict__ py, float* __restrict__ pz, int W, int H,
                                                   int D, int nreal, float a, float b)
{
    TvTile t;
    if (!tv_tile(W, H, D, nreal, t)) return;
    const size_t plane = (size_t)D * W;
    size_t k = ((size_t)t.y * D + t.z0) * W + t.x;
    const bool hx = t.x < W - 1, hy = t.y < H - 1;
    float uc = u[k];
#pragma unroll 4
    for (int z = t.z0; z < t.z1; ++z, k += W) {
        const bool hz = z < D - 1;
        const float un = hz ? u[k + W] : 0.0f;
        const float duy = hy ? u[k + plane] - uc : 0.0f;
        const float dux = hx ? u[k + 1] - uc : 0.0f;
        const float duz = hz ? un - uc : 0.0f;
        const float dd1 = 1.0f + b * fabsf(sqrtf(duy * duy + dux * dux + duz * duz));
        px[k] = (px[k] - a * dux) / dd1;
        py[k] = (py[k] - a * duy) / dd1;
        pz[k] = (pz[k] - a * duz) / dd1;
        uc = un;
    }
}

// updateU (:158-202): TomoJ's divergence takes the voxel's own component with
// a minus sign on the last column, row and slice.  With pflag, a non-finite
// result sets its row's plane flag (the caller clears the flags first).
__global__ __launch_bounds__(TJ_BLOCK) void k_tv_u(const float* __restrict__ f, float* __restrict__ u,
                                                   const float* __restrict__ px, const float* __restrict__ py,
                                                   const float* __restrict__ pz, int W, int H, int D, int nreal,
                                                   float theta, uint8_t* __restrict__ pflag)
{
    TvTile t;
    if (!tv_tile(W, H, D, nreal, t)) return;
    const size_t plane = (size_t)D * W;
    size_t k = ((size_t)t.y * D + t.z0) * W + t.x;
    float pzp = t.z0 > 0 ? pz[k - W] : 0.0f;
    bool bad = false;
#pragma unroll 4
    for (int z = t.z0; z < t.z1; ++z, k += W) {
        const float pxc = px[k], pyc = py[k], pzc = pz[k];
        float div;
        if (t.x == 0) div = pxc;
        else if (t.x == W - 1) div = -pxc;
        else div = pxc - px[k - 1];
        if (t.y == 0) div = div + pyc;
        else if (t.y == H - 1) div = div + (-pyc);
        else div = div + (pyc - py[k - plane]);
        if (z == 0) div = div + pzc;
        else if (z == D - 1) div = div + (-pzc);
        else div = div + (pzc - pzp);
        const float un = f[k] - theta * div;
        u[k] = un;
        bad |= !isfinite(un);
        pzp = pzc;
    }
    if (pflag && bad) pflag[t.y] = 1;
}

// ---------------------------------------------------------------------------
// BgART projector (BgARTVoxelProjector3D): the cross median of
// startOfIteration (medianFilter3Dcross, :123-151) and the stack statistics
// and collapse of endOfIteration (collapseFilter, :34-78).  Device layout
// [y][z][x]; TomoJ indexes a slice image as i = x + y*W.
// ---------------------------------------------------------------------------
__device__ inline int java_float_key(float v)
{
    if (v != v) return 0x7fc00000;  // Float.compare: every NaN after +inf
    const int b = __float_as_int(v);
    return b >= 0 ? b : (b ^ 0x7fffffff);
}

__global__ __launch_bounds__(TJ_BLOCK) void k_median_cross(const float* __restrict__ src, float* __restrict__ dst,
                                                           int W, int H, int D)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y, y = blockIdx.z;
    if (x >= W) return;
    const long plane = (long)W * H;
    const long i = x + (long)y * W;
    auto at = [&](int xx, int yy, int zz) { return src[((size_t)yy * D + zz) * W + xx]; };
    float v[7];
    v[0] = z > 0 ? at(x, y, z - 1) : 0.0f;
    v[1] = z < D - 1 ? at(x, y, z + 1) : 0.0f;
    v[2] = at(x, y, z);
    // index -+1 runs into the neighbouring row at the row ends
    v[3] = (i + 1 < plane) ? (x + 1 < W ? at(x + 1, y, z) : at(0, y + 1, z)) : 0.0f;
    v[4] = (i - 1 > 0) ? (x > 0 ? at(x - 1, y, z) : at(W - 1, y - 1, z)) : 0.0f;
    v[5] = (i - W > 0) ? at(x, y - 1, z) : 0.0f;
    v[6] = (i + W < plane) ? at(x, y + 1, z) : 0.0f;
    int kk[7];
#pragma unroll
    for (int a = 0; a < 7; ++a) kk[a] = java_float_key(v[a]);
    // 16-comparator sorting network on the keys; a key maps back to its float
    // (every NaN to the canonical one)
    auto cx = [&](int a, int b) {
        const int lo = min(kk[a], kk[b]), hi = max(kk[a], kk[b]);
        kk[a] = lo;
        kk[b] = hi;
    };
    cx(0, 6); cx(2, 3); cx(4, 5);
    cx(0, 2); cx(1, 4); cx(3, 6);
    cx(0, 1); cx(2, 5); cx(3, 4);
    cx(1, 2); cx(4, 6);
    cx(2, 3); cx(4, 5);
    cx(1, 2); cx(3, 4); cx(5, 6);
    const int m = kk[3];
    dst[((size_t)y * D + z) * W + x] = __int_as_float(m >= 0 ? m : (m ^ 0x7fffffff));
}

// per-block min and max over the values within +-FLT_MAX
__global__ __launch_bounds__(TJ_BLOCK) void k_stack_minmax(const float* __restrict__ vol, size_t n,
                                                           double* __restrict__ part)
{
    __shared__ double sh[TJ_BLOCK];
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const double v = vol[k];
        if (fsv_in(v)) {
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    }
    mn = block_reduce(mn, sh, MinD());
    mx = block_reduce(mx, sh, MaxD());
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

// reduces the per-block partials into part[0] (min) and part[1] (max); one block
__global__ __launch_bounds__(TJ_BLOCK) void k_stack_reduce(double* __restrict__ part, int np)
{
    __shared__ double sh[TJ_BLOCK];
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
        mn = part[2 * k] < mn ? part[2 * k] : mn;
        mx = part[2 * k + 1] > mx ? part[2 * k + 1] : mx;
    }
    mn = block_reduce(mn, sh, MinD());
    mx = block_reduce(mx, sh, MaxD());
    if (threadIdx.x == 0) {
        part[0] = mn;
        part[1] = mx;
    }
}

__global__ __launch_bounds__(TJ_BLOCK) void k_stack_hist(const float* __restrict__ vol, size_t n,
                                                         const double* __restrict__ mm, int* __restrict__ hist)
{
    __shared__ int h[256];
    for (int b = threadIdx.x; b < 256; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const double mn = mm[0], mx = mm[1];
    const double scale = 256 / (mx - mn);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const double v = vol[k];
        if (fsv_in(v) && v >= mn && v <= mx) {
            int index = java_d2i(scale * (v - mn));
            if (index >= 256) index = 255;
            atomicAdd(&h[index], 1);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 256; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ __launch_bounds__(TJ_BLOCK) void k_collapse(float* __restrict__ vol, size_t n, double th1, double th2,
                                                       float mv, int dark_bg)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const double v = vol[k];
        if (dark_bg && v < th2) vol[k] = mv;
        if (!dark_bg && v > th1) vol[k] = mv;
    }
}

// ---------------------------------------------------------------------------
// Bayesian regularisation (BayesianVoxelProjector3D): the distance volume of
// startOfIteration, convolve(rec, kernel, dist, -lambda) (:175-198), and the
// update of endOfIteration (:121-132).  Same tiles and block order as the TV
// kernels: a thread marches z down one column (x, y) and keeps the 3 x 3 x 3
// neighbourhood in registers, loading the 9 values of slice z+1 per step.
// ---------------------------------------------------------------------------
__device__ inline float bayes_vox(const float* __restrict__ v, int W, int H, int D, int x, int y, int z)
{
    // TomoReconstruction2.getPixel(int, int, int) (:241-246): 0 outside
    if (x < 0 || x >= W || y < 0 || y >= H || z < 0 || z >= D) return 0.0f;
    return v[((size_t)y * D + z) * W + x];
}

// dist = (float)(total * factor), total the double sum of the 27 float
// products value * kernel, x offset outermost and z offset innermost (the order
// of getCubicNeightborhoodWithClamp, TomoReconstruction2.java:317-339, which
// reads the unclamped neighbours).  The kernel is [1,-2,1] along each axis.
__global__ __launch_bounds__(TJ_BLOCK) void k_bayes_conv(const float* __restrict__ vol, float* __restrict__ dist,
                                                         int W, int H, int D, int nreal, double factor)
{
    TvTile t;
    if (!tv_tile(W, H, D, nreal, t)) return;
    constexpr float a[3] = {1.0f, -2.0f, 1.0f};
    float w[3][3][3];  // [i][j][k]: x offset, y offset, z offset (z-1, z, z+1)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            w[i][j][0] = bayes_vox(vol, W, H, D, t.x + i - 1, t.y + j - 1, t.z0 - 1);
            w[i][j][1] = bayes_vox(vol, W, H, D, t.x + i - 1, t.y + j - 1, t.z0);
        }
    size_t k = ((size_t)t.y * D + t.z0) * W + t.x;
    for (int z = t.z0; z < t.z1; ++z, k += W) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) w[i][j][2] = bayes_vox(vol, W, H, D, t.x + i - 1, t.y + j - 1, z + 1);
        double total = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const float prod = w[i][j][kk] * (a[i] * a[j] * a[kk]);
                    total += (double)prod;
                }
        dist[k] = (float)(total * factor);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                w[i][j][0] = w[i][j][1];
                w[i][j][1] = w[i][j][2];
            }
    }
}

// endOfIteration: dist *= (float) alpha (ImageJ FloatProcessor.multiply), rec +=
// dist (FloatBlitter ADD), then with positivity FloatBlitter MAX against 0
// (applyPositivityConstraint, TomoReconstruction2.java:405-412).  A
// non-finite result sets its row's plane flag (cleared by the caller).
__global__ __launch_bounds__(TJ_BLOCK) void k_bayes_end(float* __restrict__ vol, float* __restrict__ dist,
                                                        size_t plane, int H, float alpha, int positivity,
                                                        uint8_t* __restrict__ pflag)
{
    const int y = blockIdx.y;
    bool bad = false;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < plane; q += (size_t)gridDim.x * blockDim.x) {
        const size_t k = (size_t)y * plane + q;
        const float d = dist[k] * alpha;
        dist[k] = d;
        float v = d + vol[k];
        if (positivity && 0.0f > v) v = 0.0f;
        vol[k] = v;
        bad |= !isfinite(v);
    }
    if (bad) pflag[y] = 1;
}

// ---------------------------------------------------------------------------
// Volume layout conversion between the host ImageStack layout [z][y][x] and the
// device layout [y][z][x], for rows [y0, y0+nrows).
// ---------------------------------------------------------------------------
__global__ void k_zyx_to_yzx(const float* __restrict__ src, float* __restrict__ vol, int W, int D,
                             int nrows, int y0)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int yy = blockIdx.y, z = blockIdx.z;
    if (x >= W) return;
    vol[((size_t)(y0 + yy) * D + z) * W + x] = src[((size_t)z * nrows + yy) * W + x];
}

__global__ void k_yzx_to_zyx(const float* __restrict__ vol, float* __restrict__ dst, int W, int D,
                             int nrows, int y0)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int yy = blockIdx.y, z = blockIdx.z;
    if (x >= W) return;
    dst[((size_t)z * nrows + yy) * W + x] = vol[((size_t)(y0 + yy) * D + z) * W + x];
}

// Non-finite flag per volume plane y: blockIdx.y picks the plane, blockIdx.x
// one of gridDim.x slices of it; a slice holding a non-finite value sets the
// flag (cleared beforehand).
__global__ __launch_bounds__(TJ_BLOCK) void k_plane_flags(const float* __restrict__ vol, size_t plane,
                                                          int y0, uint8_t* __restrict__ pflag)
{
    __shared__ unsigned sh[TJ_BLOCK];
    const int y = y0 + blockIdx.y;
    const float* p = vol + (size_t)y * plane;
    const size_t per = (plane + gridDim.x - 1) / gridDim.x;
    const size_t k0 = per * blockIdx.x, k1 = min(plane, k0 + per);
    unsigned bad = 0;
#pragma unroll 4
    for (size_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) bad |= isfinite(p[k]) ? 0u : 1u;
    bad = block_reduce(bad, sh, OrU());
    if (threadIdx.x == 0 && bad) pflag[y] = 1;
}

// ---------------------------------------------------------------------------
// Synthetic tilt series: analytic FP64 line integrals of an ellipsoid phantom
// in TomoJ's geometry, then Gaussian noise.
// ---------------------------------------------------------------------------
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// pixel (i,j) of image p: line u*(i-pc0) + v*(j-pc1) + t*w, u,v,w = rows of E_p.
__global__ __launch_bounds__(TJ_BLOCK) void k_phantom_project(
    const double* __restrict__ ell, int k, const TjProj* __restrict__ pp, int pw, int ph, double pc0,
    double pc1, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, p = blockIdx.z;
    if (i >= pw) return;
    const double* E = pp[p].e;
    const double a = i - pc0, b = j - pc1;
    const double pt0 = E[0] * a + E[3] * b, pt1 = E[1] * a + E[4] * b, pt2 = E[2] * a + E[5] * b;
    const double d0 = E[6], d1 = E[7], d2 = E[8];
    double sum = 0.0;
    for (int e = 0; e < k; ++e) {
        const double* o = ell + 16 * e;
        const double* R = o + 7;
        const double q0 = pt0 - o[1], q1 = pt1 - o[2], q2 = pt2 - o[3];
        double A = 0, B = 0, C = 0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const double qa = R[3 * ax] * q0 + R[3 * ax + 1] * q1 + R[3 * ax + 2] * q2;
            const double da = R[3 * ax] * d0 + R[3 * ax + 1] * d1 + R[3 * ax + 2] * d2;
            const double inv = 1.0 / (o[4 + ax] * o[4 + ax]);
            A += da * da * inv;
            B += 2 * da * qa * inv;
            C += qa * qa * inv;
        }
        C -= 1;
        const double disc = B * B - 4 * A * C;
        if (disc > 0) sum += o[0] * sqrt(disc) / A;
    }
    out[((size_t)p * ph + j) * pw + i] = sum;
}

__global__ __launch_bounds__(TJ_BLOCK) void k_max_partial(const double* __restrict__ a, size_t n,
                                                          double* __restrict__ part)
{
    __shared__ double sh[TJ_BLOCK];
    double m = -INFINITY;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, a[k]);
    m = block_reduce(m, sh, MaxD());
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// value n: a[n] + sigma * N(0,1), the normal deviate from mix64(seed ^ 2n) and
// mix64(seed ^ (2n+1)) through Box-Muller.
__global__ __launch_bounds__(TJ_BLOCK) void k_add_noise(const double* __restrict__ a, size_t n,
                                                        uint64_t seed, const double* __restrict__ sigma_src,
                                                        int nsig, double rel, float* __restrict__ out)
{
    __shared__ double s_sigma;
    if (threadIdx.x == 0) {
        double m = -INFINITY;
        for (int t = 0; t < nsig; ++t) m = fmax(m, sigma_src[t]);
        s_sigma = rel * m;
    }
    __syncthreads();
    const double sigma = s_sigma;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const double u1 = (double)((mix64(seed ^ (uint64_t)(2 * k)) >> 11) + 1) * 0x1.0p-53;
        const double u2 = (double)(mix64(seed ^ (uint64_t)(2 * k + 1)) >> 11) * 0x1.0p-53;
        const double gz = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
        out[k] = (float)(a[k] + sigma * gz);
    }
}

// Plain streaming copy (16-byte vectors), used to calibrate HBM byte counters.
// VoxelProjector3D.convertDiffToErrorImage (VoxelProjector3D.java:315-324):
// every difference pixel squared in FP32, NaN left as it is.
__global__ __launch_bounds__(TJ_BLOCK) void k_square_diffs(float* __restrict__ d, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * TJ_BLOCK + threadIdx.x; k < n; k += (size_t)gridDim.x * TJ_BLOCK) {
        const float f = d[k];
        if (!isnan(f)) d[k] = f * f;
    }
}

__global__ __launch_bounds__(TJ_BLOCK) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   size_t n4)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

}  // namespace tj

// ---------------------------------------------------------------------------
// Launch wrappers (C++ linkage, used by tj_api.hip).
// ---------------------------------------------------------------------------
namespace tjk {
using namespace tj;

static int g_max_lds = 65536;

void set_max_lds(int bytes) { g_max_lds = bytes; }

template <int G, bool FUSE = false, int GB = 1>
static hipError_t launch_fp_lds(hipStream_t st, const TjGeomDev& g, const float* vol, const TjProj* pp,
                                const int* slot_proj, int nslot, int start_row, const float* norm_tab, int zc,
                                float* proj, float* norm, TjFpFuse fu, bool pf)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    const int W1 = g.vw + 1;
    const int per = W1 * (int)sizeof(float4);
    // the fused form's reduction arrays are static LDS beside the staged slices
    const int fixed = FUSE ? (TJ_BLOCK * GB / 64) * G * (int)(sizeof(double) + 4 * sizeof(unsigned)) : 0;
    // double-buffered form: two slabs, each thread stages at most TJ_PFQ cells
    // of a chunk; used where two or more slices fit
    if (pf) {
        const int zpf = min(((g_max_lds - fixed) / 2 - 2 * (int)sizeof(float4)) / per, TJ_PFQ * TJ_BLOCK * GB / W1);
        pf = zpf >= 2 && g.vd >= 2;
        if (pf && (zc <= 0 || zc > zpf)) zc = zpf;
    }
    if (!pf) {
        const int zmax = (g_max_lds - 2 * (int)sizeof(float4) - fixed) / per;
        if (zmax < 1) return hipErrorInvalidValue;
        if (zc <= 0 || zc > zmax) zc = zmax;
    }
    if (zc > g.vd) zc = g.vd;
    const size_t lds = (size_t)(pf ? 2 : 1) * (zc * W1 + 2) * sizeof(float4);
    const int nrb = (g.ph + 3) / 4;
    const int ntile = (g.pw + TJ_BLOCK - 1) / TJ_BLOCK;
    const int ngrp = (nslot + G * GB - 1) / (G * GB);
    const int nsub = ntile * ngrp;
    const int nblk = ((nrb + 7) / 8) * 8 * nsub;
    auto go = [&](auto kern) -> hipError_t {
        if (lds + fixed > 65536) {
            hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        kern<<<nblk, TJ_BLOCK * GB, lds, st>>>(g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, nrb, nsub,
                                               ntile, proj, norm, fu);
        return hipGetLastError();
    };
    return pf ? go(k_fp_lds<G, FUSE, GB, true>) : go(k_fp_lds<G, FUSE, GB, false>);
}

// The LDS forward kernel with gb groups of G projections per block (1, 2 or 4;
// G = 8 takes 1).  gb = 0 picks 4 with the double-buffered slab (pf; one block
// per CU hides its own load latency, and 1024 threads stage up to 7 slices of
// a 512-column slab); with a single buffer, a launch of at most 4 groups (an
// OS-SART subset) in one block per row group and column tile, so the slab is
// read once, and more groups (SIRT) in pairs, which keeps two blocks per CU to
// overlap one block's staging with the other's rays.
template <int G, bool FUSE = false>
static hipError_t launch_fp_lds_gb(int gb, hipStream_t st, const TjGeomDev& g, const float* vol, const TjProj* pp,
                                   const int* slot_proj, int nslot, int start_row, const float* norm_tab, int zc,
                                   float* proj, float* norm, TjFpFuse fu, bool pf)
{
    if (gb <= 0) gb = (pf || (nslot + G - 1) / G <= 4) ? 4 : 2;
    if constexpr (G <= 4) {
        if (gb >= 4)
            return launch_fp_lds<G, FUSE, 4>(st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, proj, norm, fu,
                                             pf);
        if (gb == 2)
            return launch_fp_lds<G, FUSE, 2>(st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, proj, norm, fu,
                                             pf);
    }
    return launch_fp_lds<G, FUSE, 1>(st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, proj, norm, fu, pf);
}

hipError_t norm_table(hipStream_t st, const TjGeomDev& g, const TjProj* pp, int nproj, float* tab)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nproj <= 0) return hipSuccess;
    k_norm_table<<<dim3((g.pw + TJ_BLOCK - 1) / TJ_BLOCK, nproj), TJ_BLOCK, 0, st>>>(g, pp, tab);
    return hipGetLastError();
}

// kernel: 1 generic, 2 axis (global loads), 3 axis with LDS staging
hipError_t fp(hipStream_t st, const TjGeomDev& g, const float* vol, const TjProj* pp, const int* slot_proj,
              int nslot, int start_row, int kernel, const uint8_t* pflag, const float* norm_tab, float* proj,
              float* norm, int rows, int rpt, int zc, int gb, bool pf)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nslot <= 0) return hipSuccess;
    const int bx = (g.pw + TJ_BLOCK - 1) / TJ_BLOCK;
    if (kernel == 1) {
        k_fp_generic<<<dim3(bx, g.ph, nslot), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, start_row, proj, norm);
        return hipGetLastError();
    }
    if (kernel == 3) {
        if (rpt <= 2)
            return launch_fp_lds_gb<2>(gb, st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, proj, norm,
                                          TjFpFuse{}, pf);
        if (rpt <= 4)
            return launch_fp_lds_gb<4>(gb, st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, proj, norm,
                                          TjFpFuse{}, pf);
        return launch_fp_lds_gb<8>(gb, st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, proj, norm,
                                          TjFpFuse{}, pf);
    }
    switch (rows) {
    case 1:
        k_fp_axis<1><<<dim3(bx, g.ph, nslot), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, start_row, pflag, proj, norm);
        break;
    case 2:
        k_fp_axis<2><<<dim3(bx, (g.ph + 1) / 2, nslot), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, start_row, pflag, proj, norm);
        break;
    case 8:
        k_fp_axis<8><<<dim3(bx, (g.ph + 7) / 8, nslot), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, start_row, pflag, proj, norm);
        break;
    default:
        k_fp_axis<4><<<dim3(bx, (g.ph + 3) / 4, nslot), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, start_row, pflag, proj, norm);
        break;
    }
    return hipGetLastError();
}

hipError_t norm_max_fill(hipStream_t st, int npx, int nslot, float* norm)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nslot <= 0) return hipSuccess;
    k_norm_max_fill<<<nslot, TJ_BLOCK, 0, st>>>(npx, norm);
    return hipGetLastError();
}

hipError_t residual(hipStream_t st, int pw, int ph, int nslot, const float* exps, const int* exp_index,
                    const float* proj, const float* norm, double factor, float* diff, double* row_err,
                    uint8_t* rowflag, double* err)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nslot <= 0) return hipSuccess;
    k_residual<<<dim3(ph, nslot), TJ_BLOCK, 0, st>>>(pw, ph, exps, exp_index, proj, norm, factor, diff, row_err, rowflag);
    k_err_sum<<<nslot, TJ_BLOCK, 0, st>>>(ph, row_err, err);
    return hipGetLastError();
}

// Per slot: the sum of the error partials of its row blocks and column tiles
// (in that order), and each row's non-finite flag over its column tiles.
// clear (or null): an int zeroed on the way (the backprojection's row flag,
// so that the next launch needs no memset of its own)
__global__ __launch_bounds__(TJ_BLOCK) void k_err_fold(int ph, int ntile, const double* __restrict__ rerr,
                                                       const uint8_t* __restrict__ rbad,
                                                       uint8_t* __restrict__ rowflag, double* __restrict__ err,
                                                       int* __restrict__ clear)
{
    __shared__ double sh[TJ_BLOCK];
    const int s = blockIdx.x;
    if (clear && s == 0 && threadIdx.x == 0) *clear = 0;
    const int nq = ((ph + 3) / 4) * ntile;
    double acc = 0.0;
    for (int q = threadIdx.x; q < nq; q += blockDim.x) acc += rerr[(size_t)s * nq + q];
    for (int j = threadIdx.x; j < ph; j += blockDim.x) {
        const size_t q = ((size_t)s * ph + j) * ntile;
        uint8_t b = 0;
        for (int t = 0; t < ntile; ++t) b |= rbad[q + t];
        rowflag[(size_t)s * ph + j] = b;
    }
    acc = block_reduce(acc, sh, AddD());
    if (threadIdx.x == 0) err[s] = acc;
}

// The forward projection of the LDS kernel fused with the residual (the
// projectAndDifference of one subset): diff, rowflag and the per-slot errors as
// fp + residual give them (the errors summed in another order), and proj and
// norm too when they are given.  rerr holds nslot x ceil(ph / 4) x ceil(pw / 256)
// entries, rbad nslot x ph x ceil(pw / 256).  `mid` (or null) is recorded
// between the forward kernel and the fold.
hipError_t fp_residual(hipStream_t st, const TjGeomDev& g, const float* vol, const TjProj* pp, const int* slot_proj,
                       int nslot, int start_row, const float* norm_tab, int rpt, int zc, const float* exps,
                       double factor, float* diff, double* rerr, uint8_t* rbad, uint8_t* rowflag, double* err,
                       hipEvent_t mid, int gb, int* clear, bool pf, float* proj, float* norm)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nslot <= 0) return hipSuccess;
    const TjFpFuse fu{exps, factor, diff, rerr, rbad, proj, norm};
    hipError_t e;
    if (rpt <= 2)
        e = launch_fp_lds_gb<2, true>(gb, st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, nullptr, nullptr,
                                      fu, pf);
    else if (rpt <= 4)
        e = launch_fp_lds_gb<4, true>(gb, st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, nullptr, nullptr,
                                      fu, pf);
    else
        e = launch_fp_lds_gb<8, true>(gb, st, g, vol, pp, slot_proj, nslot, start_row, norm_tab, zc, nullptr, nullptr,
                                      fu, pf);
    if (e != hipSuccess) return e;
    if (mid && (e = hipEventRecord(mid, st)) != hipSuccess) return e;
    k_err_fold<<<nslot, TJ_BLOCK, 0, st>>>(g.ph, (g.pw + TJ_BLOCK - 1) / TJ_BLOCK, rerr, rbad, rowflag, err, clear);
    return hipGetLastError();
}

hipError_t row_flags(hipStream_t st, int pw, int ph, int nslot, const float* images, const int* image_index,
                     uint8_t* rowflag)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nslot <= 0) return hipSuccess;
    k_row_flags<<<dim3(ph, nslot), TJ_BLOCK, 0, st>>>(pw, ph, images, image_index, rowflag);
    return hipGetLastError();
}

template <int R, int KC, int G>
static hipError_t launch_bp_lds(hipStream_t st, const TjGeomDev& g, float* vol, const TjProj* pp,
                                const int* slot_proj, int nslot, const float* images, const int* image_index,
                                const uint8_t* rowflag, TjBpRow* rows, int start_y, int end_y, int finite_volume,
                                double2* pairs, hipEvent_t mid, int* slow, int fast)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    const size_t npx = (size_t)g.pw * g.ph;
    const bool fast_on = slow && fast > 0 && finite_volume && !g.positivity;
    const int nrb = (g.vh + TJ_BLOCK - 1) / TJ_BLOCK;
    const unsigned npb = (unsigned)std::min<size_t>((npx + TJ_BLOCK - 1) / TJ_BLOCK, 1024);
    k_bp_prep<<<dim3(nrb + npb, nslot), TJ_BLOCK, 0, st>>>(g, pp, slot_proj, nslot, rowflag, rows, slow, nrb, images,
                                                           image_index, pairs);
    if (mid) {
        const hipError_t e = hipEventRecord(mid, st);
        if (e != hipSuccess) return e;
    }
    const int ny = end_y - start_y;
    // the fast kernel (finite volume, no positivity; it leaves the launch to
    // the general kernel when k_bp_rows flagged a row it cannot take)
    const int* run_if = nullptr;
    if (fast_on) {
        run_if = slow;
        auto grid_of = [&](int r, int kc) {
            return (unsigned)(((g.vw + TJ_BLOCK - 1) / TJ_BLOCK) * ((ny + r - 1) / r) * ((g.vd + kc - 1) / kc));
        };
#define TJ_BPF(RR, KK, WW) k_bp_fast<RR, KK, WW><<<grid_of(RR, KK), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, rows, start_y, end_y, pairs, slow)
        switch (fast) {
        case 1: TJ_BPF(4, 8, 4); break;
        case 2: TJ_BPF(4, 8, 3); break;
        case 3: TJ_BPF(4, 4, 4); break;
        case 4: TJ_BPF(2, 8, 4); break;
        case 5: TJ_BPF(4, 16, 2); break;
        case 6: TJ_BPF(2, 16, 3); break;
        default: TJ_BPF(4, 4, 5); break;
        }
#undef TJ_BPF
    }
    const unsigned grid = (unsigned)(((g.vw + TJ_BLOCK - 1) / TJ_BLOCK) * ((ny + R - 1) / R) * ((g.vd + KC - 1) / KC));
    if (g.positivity)
        k_bp_lds<R, KC, G, true><<<grid, TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, rows, start_y, end_y, finite_volume, pairs, run_if);
    else
        k_bp_lds<R, KC, G, false><<<grid, TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, rows, start_y, end_y, finite_volume, pairs, run_if);
    return hipGetLastError();
}

// kernel: 1 generic, 2 axis (global loads), 3 axis with LDS staging (scale 1)
hipError_t bp(hipStream_t st, const TjGeomDev& g, float* vol, const TjProj* pp, const int* slot_proj, int nslot,
              const float* images, const int* image_index, const uint8_t* rowflag, TjBpRow* rows, int start_y,
              int end_y, int kernel, int rows_per_thread, int kchunk, int group, int finite_volume,
              double2* pairs, hipEvent_t mid, int* slow, int fast, int slow_cleared)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nslot <= 0 || end_y <= start_y) return hipSuccess;
    const int bx = (g.vw + TJ_BLOCK - 1) / TJ_BLOCK;
    const int ny = end_y - start_y;
    // `mid` (or null) is recorded after the preparation kernels (row table,
    // pair images), before the backprojection kernel itself
    if (kernel == 1) {
        if (mid) {
            const hipError_t e = hipEventRecord(mid, st);
            if (e != hipSuccess) return e;
        }
        if (g.positivity)
            k_bp_generic<true><<<dim3(bx, ny, g.vd), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, start_y);
        else
            k_bp_generic<false><<<dim3(bx, ny, g.vd), TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, start_y);
        return hipGetLastError();
    }
    // `slow` (two device ints, or null): slow[0] set by the row table when a
    // row needs the general kernel, slow[1] counts the launches the fast kernel
    // took; only the LDS path with a fast shape uses them.  slow[0] is cleared
    // here unless the caller's previous launch on the stream (k_err_fold)
    // cleared it.
    int* slow_flag = (kernel == 3 && fast > 0 && finite_volume && !g.positivity) ? slow : nullptr;
    if (slow_flag && !slow_cleared) {
        const hipError_t e = hipMemsetAsync(slow_flag, 0, sizeof(int), st);
        if (e != hipSuccess) return e;
    }
    if (kernel == 3) {
#define TJ_BPL(RR, KK, GG) return launch_bp_lds<RR, KK, GG>(st, g, vol, pp, slot_proj, nslot, images, image_index, rowflag, rows, start_y, end_y, finite_volume, pairs, mid, slow_flag, fast)
        if (rows_per_thread == 2) {
            if (kchunk >= 16) { if (group >= 4) TJ_BPL(2, 16, 4); TJ_BPL(2, 16, 2); }
            if (group >= 8) TJ_BPL(2, 8, 8);
            if (group >= 4) TJ_BPL(2, 8, 4);
            TJ_BPL(2, 8, 2);
        }
        if (rows_per_thread == 8) TJ_BPL(8, 4, 1);
        if (kchunk <= 4) TJ_BPL(4, 4, 2);
        if (group >= 4) TJ_BPL(4, 8, 4);
        if (group >= 3) TJ_BPL(4, 8, 3);
        TJ_BPL(4, 8, 2);
#undef TJ_BPL
    }
    k_bp_rows<<<dim3((g.vh + 255) / 256, nslot), 256, 0, st>>>(g, pp, slot_proj, nslot, rowflag, rows, nullptr);
    if (mid) {
        const hipError_t e = hipEventRecord(mid, st);
        if (e != hipSuccess) return e;
    }
    if (kchunk <= 0) kchunk = 8;
    const int bz = (g.vd + kchunk - 1) / kchunk;
    const bool s1 = (g.scale == 1.0);
    const bool pos = g.positivity != 0;
#define TJ_BP_LAUNCH(RR)                                                                                     \
    {                                                                                                        \
        dim3 grid(bx, (ny + RR - 1) / RR, bz);                                                               \
        if (pos && s1) k_bp_axis<RR, true, true><<<grid, TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, rows, start_y, end_y, kchunk); \
        else if (pos) k_bp_axis<RR, true, false><<<grid, TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, rows, start_y, end_y, kchunk); \
        else if (s1) k_bp_axis<RR, false, true><<<grid, TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, rows, start_y, end_y, kchunk); \
        else k_bp_axis<RR, false, false><<<grid, TJ_BLOCK, 0, st>>>(g, vol, pp, slot_proj, nslot, images, image_index, rows, start_y, end_y, kchunk); \
    }
    switch (rows_per_thread) {
    case 1: TJ_BP_LAUNCH(1) break;
    case 2: TJ_BP_LAUNCH(2) break;
    case 8: TJ_BP_LAUNCH(8) break;
    default: TJ_BP_LAUNCH(4) break;
    }
#undef TJ_BP_LAUNCH
    return hipGetLastError();
}

hipError_t roi_stats(hipStream_t st, const float* imgs, int pw, int ph, int n, const int* roi, double* stats)
{
    (void)hipGetLastError();
    if (n <= 0) return hipSuccess;
    k_roi_stats<<<n, TJ_BLOCK, 0, st>>>(imgs, pw, ph, roi, stats);
    return hipGetLastError();
}

hipError_t poly_stats(hipStream_t st, const float* imgs, int pw, int ph, int n, const int* poly, double* stats)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n <= 0) return hipSuccess;
    k_poly_stats<<<n, TJ_BLOCK, 0, st>>>(imgs, pw, ph, poly, stats);
    return hipGetLastError();
}

hipError_t affine(hipStream_t st, const float* src, float* dst, int w, int h, int n, const double* inv,
                  const int* ident, const double* stats, int normalize, int fill)
{
    (void)hipGetLastError();
    if (n <= 0) return hipSuccess;
    k_affine<<<dim3((w + TJ_BLOCK - 1) / TJ_BLOCK, h, n), TJ_BLOCK, 0, st>>>(src, dst, w, h, inv, ident, stats,
                                                                              normalize, fill);
    return hipGetLastError();
}

hipError_t normalize(hipStream_t st, float* imgs, int pw, int ph, int n, const int* roi, double* stats)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n <= 0) return hipSuccess;
    const size_t npx = (size_t)pw * ph;
    k_roi_stats<<<n, TJ_BLOCK, 0, st>>>(imgs, pw, ph, roi, stats);
    k_normalize<<<dim3((unsigned)std::min<size_t>((npx + TJ_BLOCK - 1) / TJ_BLOCK, 1024), n), TJ_BLOCK, 0, st>>>(
        imgs, npx, stats);
    return hipGetLastError();
}

hipError_t fista(hipStream_t st, float* vol, const float* prev, size_t n, double factor)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    const int blocks = (int)std::min<size_t>((n + TJ_BLOCK - 1) / TJ_BLOCK, 8192);
    k_fista<<<blocks, TJ_BLOCK, 0, st>>>(vol, prev, n, factor);
    return hipGetLastError();
}

// TiltSeries.applyThreshold on n resident images (in place); lv: 4 doubles
// per image, q: a byte per pixel, changed: one int (device scratch).
hipError_t threshold(hipStream_t st, float* imgs, int w, int h, int n, double thr, double hys, int fill, double* lv,
                     uint8_t* q, int* changed)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n <= 0 || w <= 0 || h <= 0) return hipSuccess;
    if (n > 65535) return hipErrorInvalidValue;
    const size_t len = (size_t)w * h;
    k_thr_levels<<<n, TJ_BLOCK, 0, st>>>(imgs, w, h, thr, hys, fill, lv);
    const unsigned bx = (unsigned)std::min<size_t>((len + TJ_BLOCK - 1) / TJ_BLOCK, 1024);
    k_thr_apply<<<dim3(bx, n), TJ_BLOCK, 0, st>>>(imgs, len, lv, q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // grow until a pass changes nothing; four passes per check of the flag
    int hc = 1;
    for (size_t pass = 0; hc && pass < len + 4; pass += 4) {
        if ((e = hipMemsetAsync(changed, 0, sizeof(int), st)) != hipSuccess) return e;
        for (int k = 0; k < 4; ++k) k_thr_grow<<<dim3(bx, n), TJ_BLOCK, 0, st>>>(imgs, len, w, lv, q, changed);
        if ((e = hipMemcpyAsync(&hc, changed, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t median_cross(hipStream_t st, const float* src, float* dst, int W, int H, int D)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if ((size_t)W * H * D == 0) return hipSuccess;
    // blocks in (x tile, z, y) order: the rows and slices a block reads around
    // its own are read by blocks running at the same time.  Handing each XCD a
    // contiguous eighth of this order measured slower (0.34 against 0.27 ms at
    // 512 x 512 x 256).
    if (H > 65535 || D > 65535) return hipErrorInvalidValue;
    k_median_cross<<<dim3((W + TJ_BLOCK - 1) / TJ_BLOCK, D, H), TJ_BLOCK, 0, st>>>(src, dst, W, H, D);
    return hipGetLastError();
}

// StackStatistics' min, max (mm[0..1]) and 256-bin histogram of the volume;
// mm needs 2 x 4096 doubles, hist 256 ints.  Both passes run 4096 blocks; each
// histogram block merges its non-empty bins into the global histogram.
hipError_t stack_stats(hipStream_t st, const float* vol, size_t n, double* mm, int* hist)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n == 0) return hipErrorInvalidValue;
    const int blocks = (int)std::min<size_t>((n + TJ_BLOCK - 1) / TJ_BLOCK, 4096);
    hipError_t e = hipMemsetAsync(hist, 0, 256 * sizeof(int), st);
    if (e != hipSuccess) return e;
    k_stack_minmax<<<blocks, TJ_BLOCK, 0, st>>>(vol, n, mm);
    k_stack_reduce<<<1, TJ_BLOCK, 0, st>>>(mm, blocks);
    k_stack_hist<<<blocks, TJ_BLOCK, 0, st>>>(vol, n, mm, hist);
    return hipGetLastError();
}

hipError_t collapse(hipStream_t st, float* vol, size_t n, double th1, double th2, float mv, int dark_bg)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + TJ_BLOCK - 1) / TJ_BLOCK, 8192);
    k_collapse<<<blocks, TJ_BLOCK, 0, st>>>(vol, n, th1, th2, mv, dark_bg);
    return hipGetLastError();
}

// denoise3D (:84-113): f keeps the input, u (the volume) starts equal to it,
// p starts at zero.
hipError_t tv_denoise(hipStream_t st, float* vol, float* f, float* p, int W, int H, int D, float theta, float g,
                      float dt, int iterations, uint8_t* pflag)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    const size_t n = (size_t)W * H * D;
    if (n == 0 || iterations <= 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(f, vol, n * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(p, 0, 3 * n * sizeof(float), st);
    if (e != hipSuccess) return e;
    const float a = dt / theta;
    const float b = dt / theta / g;
    const int nxt = (W + TJ_BLOCK - 1) / TJ_BLOCK;
    const long nreal = (long)nxt * H * ((D + TJ_TV_ZR - 1) / TJ_TV_ZR);
    if (nreal > 0x7ffffff0L) return hipErrorInvalidValue;
    const int nb = (int)((nreal + 7) & ~7L);
    if (pflag && (e = hipMemsetAsync(pflag, 0, H, st)) != hipSuccess) return e;
    for (int it = 0; it < iterations; ++it) {
        k_tv_p<<<nb, TJ_BLOCK, 0, st>>>(vol, p, p + n, p + 2 * n, W, H, D, (int)nreal, a, b);
        k_tv_u<<<nb, TJ_BLOCK, 0, st>>>(f, vol, p, p + n, p + 2 * n, W, H, D, (int)nreal, theta,
                                        it == iterations - 1 ? pflag : nullptr);
    }
    return hipGetLastError();
}

hipError_t bayes_conv(hipStream_t st, const float* vol, float* dist, int W, int H, int D, double factor)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    const int nxt = (W + TJ_BLOCK - 1) / TJ_BLOCK;
    const long nreal = (long)nxt * H * ((D + TJ_TV_ZR - 1) / TJ_TV_ZR);
    if (nreal <= 0) return hipSuccess;
    if (nreal > 0x7ffffff0L) return hipErrorInvalidValue;
    k_bayes_conv<<<(int)((nreal + 7) & ~7L), TJ_BLOCK, 0, st>>>(vol, dist, W, H, D, (int)nreal, factor);
    return hipGetLastError();
}

hipError_t bayes_end(hipStream_t st, float* vol, float* dist, size_t plane, int H, float alpha, int positivity,
                     uint8_t* pflag)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (H <= 0 || plane == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(pflag, 0, H, st);
    if (e != hipSuccess) return e;
    const unsigned bx = (unsigned)std::min<size_t>(64, (plane + 4 * TJ_BLOCK - 1) / (4 * TJ_BLOCK));
    k_bayes_end<<<dim3(bx, H), TJ_BLOCK, 0, st>>>(vol, dist, plane, H, alpha, positivity, pflag);
    return hipGetLastError();
}

hipError_t square_diffs(hipStream_t st, float* d, size_t n)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + TJ_BLOCK - 1) / TJ_BLOCK, 4096);
    k_square_diffs<<<blocks, TJ_BLOCK, 0, st>>>(d, n);
    return hipGetLastError();
}

hipError_t copy(hipStream_t st, const void* src, void* dst, size_t bytes)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    k_copy<<<2048, TJ_BLOCK, 0, st>>>((const float4*)src, (float4*)dst, bytes / 16);
    return hipGetLastError();
}

hipError_t zyx_to_yzx(hipStream_t st, const float* src, float* vol, int W, int D, int nrows, int y0)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nrows <= 0) return hipSuccess;
    k_zyx_to_yzx<<<dim3((W + 255) / 256, nrows, D), 256, 0, st>>>(src, vol, W, D, nrows, y0);
    return hipGetLastError();
}

hipError_t yzx_to_zyx(hipStream_t st, const float* vol, float* dst, int W, int D, int nrows, int y0)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nrows <= 0) return hipSuccess;
    k_yzx_to_zyx<<<dim3((W + 255) / 256, nrows, D), 256, 0, st>>>(vol, dst, W, D, nrows, y0);
    return hipGetLastError();
}

hipError_t plane_flags(hipStream_t st, const float* vol, size_t plane, int y0, int nrows, uint8_t* pflag)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (nrows <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(pflag + y0, 0, nrows, st);
    if (e != hipSuccess) return e;
    const unsigned split = (unsigned)std::min<size_t>(16, (plane + 4095) / 4096);
    k_plane_flags<<<dim3(split, nrows), TJ_BLOCK, 0, st>>>(vol, plane, y0, pflag);
    return hipGetLastError();
}

hipError_t phantom_project(hipStream_t st, const double* ell, int k, const TjProj* pp, int n, int pw, int ph,
                           double pc0, double pc1, double* out)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    k_phantom_project<<<dim3((pw + TJ_BLOCK - 1) / TJ_BLOCK, ph, n), TJ_BLOCK, 0, st>>>(ell, k, pp, pw, ph, pc0, pc1, out);
    return hipGetLastError();
}

hipError_t add_noise(hipStream_t st, const double* a, size_t n, uint64_t seed, double rel, double* part,
                     int npart, float* out)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    k_max_partial<<<npart, TJ_BLOCK, 0, st>>>(a, n, part);
    k_add_noise<<<1024, TJ_BLOCK, 0, st>>>(a, n, seed, part, npart, rel, out);
    return hipGetLastError();
}

}  // namespace tjk
