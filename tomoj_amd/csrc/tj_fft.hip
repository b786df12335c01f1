// tj_fft.hip -- weighting filters for weighted backprojection.
//
// FFTWeighting.weighting (filters/FFTWeighting.java:87-112) filters each image
// with real(ifft2(mask * fft2(img))), the inverse scaled by 1/(W*H).  Here the
// 2D transform is done as row FFTs, a tiled transpose, row FFTs of the
// transposed data (the column transforms, with the mask multiplied in), the
// inverse row FFTs, a transpose back and the inverse row FFTs writing the real
// part.  Row transforms of power-of-two length run as a radix-2 Stockham FFT in
// LDS (one block per row); other lengths use a direct DFT with FP64 sums.
// Twiddles are computed in FP64 and rounded once to FP32.
#include "tj_device.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace tjf {

constexpr int kMaxBatch = 16;   // images transformed per pass (bounds the workspace)

__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// tw[m] = exp(-2 pi i m / n), m < n.
__global__ void k_twiddles(int n, float2* __restrict__ tw)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
}

// Input/output forms of a row transform.
enum RowMode {
    kRealIn = 0,      // in: float image rows (image index table), out: complex
    kComplex = 1,     // complex -> complex
    kComplexMask = 2, // complex -> complex, then multiply element (row r, col c) by mask[c][r]
    kRealOut = 3      // complex -> real part * scale into float images
};

template <int MODE>
__device__ inline float2 load_in(const void* in, const int* idx, int rows_per_img, int n, int row, int k, size_t img_px)
{
    if (MODE == kRealIn) {
        const int img = row / rows_per_img, r = row % rows_per_img;
        const float* p = (const float*)in + (size_t)idx[img] * img_px + (size_t)r * n;
        return make_float2(p[k], 0.0f);
    } else {
        return ((const float2*)in)[(size_t)row * n + k];
    }
}

template <int MODE>
__device__ inline void store_out(void* out, const float* mask, int rows_per_img, int n, int row, int k, float2 v,
                                 float scale)
{
    if (MODE == kRealOut) {
        ((float*)out)[(size_t)row * n + k] = v.x * scale;
    } else if (MODE == kComplexMask) {
        const int r = row % rows_per_img;  // transposed image: row r = image column, k = image row
        const float m = mask[(size_t)k * rows_per_img + r];
        ((float2*)out)[(size_t)row * n + k] = make_float2(v.x * m, v.y * m);
    } else {
        ((float2*)out)[(size_t)row * n + k] = v;
    }
}

// Radix-2 Stockham FFT of one row of length n = 2^logn per block.  inverse:
// conjugate twiddles (no scaling).  Dynamic LDS: 2*n float2.
template <int MODE>
__global__ __launch_bounds__(256) void k_fft_row(const void* __restrict__ in, const int* __restrict__ idx,
                                                 void* __restrict__ out, const float* __restrict__ mask,
                                                 const float2* __restrict__ tw, int n, int logn, int rows_per_img,
                                                 size_t img_px, int inverse, float scale)
{
    extern __shared__ float2 lds[];
    float2* a = lds;
    float2* b = lds + n;
    const int row = blockIdx.x;
    for (int k = threadIdx.x; k < n; k += blockDim.x) a[k] = load_in<MODE>(in, idx, rows_per_img, n, row, k, img_px);
    __syncthreads();
    const int half = n >> 1;
    for (int s = 0; s < logn; ++s) {
        const int ns = 1 << s;
        for (int j = threadIdx.x; j < half; j += blockDim.x) {
            const int kk = j & (ns - 1);
            float2 w = tw[kk * (half >> s)];
            if (inverse) w.y = -w.y;
            const float2 v0 = a[j];
            const float2 v1 = cmul(a[j + half], w);
            const int d = ((j >> s) << (s + 1)) + kk;
            b[d] = make_float2(v0.x + v1.x, v0.y + v1.y);
            b[d + ns] = make_float2(v0.x - v1.x, v0.y - v1.y);
        }
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) store_out<MODE>(out, mask, rows_per_img, n, row, k, a[k], scale);
}

// Direct DFT of one row (any n): out[k] = sum_t x[t] exp(-+2 pi i k t / n), FP64 sums.
template <int MODE>
__global__ __launch_bounds__(256) void k_dft_row(const void* __restrict__ in, const int* __restrict__ idx,
                                                 void* __restrict__ out, const float* __restrict__ mask,
                                                 const float2* __restrict__ tw, int n, int rows_per_img,
                                                 size_t img_px, int inverse, float scale)
{
    const int row = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < n; ++t) {
        const float2 x = load_in<MODE>(in, idx, rows_per_img, n, row, t, img_px);
        float2 w = tw[(int)(((long long)k * t) % n)];
        if (inverse) w.y = -w.y;
        sr += (double)x.x * w.x - (double)x.y * w.y;
        si += (double)x.x * w.y + (double)x.y * w.x;
    }
    store_out<MODE>(out, mask, rows_per_img, n, row, k, make_float2((float)sr, (float)si), scale);
}

// Batched complex transpose: nb images of rows x cols -> cols x rows.
__global__ void k_transpose(const float2* __restrict__ in, float2* __restrict__ out, int rows, int cols)
{
    __shared__ float2 tile[32][33];
    const size_t img = (size_t)blockIdx.z * rows * cols;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
        const int r = r0 + dy, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[dy][threadIdx.x] = in[img + (size_t)r * cols + c];
    }
    __syncthreads();
    for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
        const int c = c0 + dy, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[img + (size_t)c * rows + r] = tile[threadIdx.x][dy];
    }
}

// FFTWeighting.createWeightingMask (FFTWeighting.java:31-65).  sint/cosp/sinp are
// sin(tilt), cos(axis), sin(axis) per image, computed on the host.
__global__ __launch_bounds__(256) void k_radermacher(int sx, int sy, double diameter, const double* __restrict__ sint,
                                                     const double* __restrict__ cosp, const double* __restrict__ sinp,
                                                     int n, float* __restrict__ mask)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= sx) return;
    const double cx = sx / 2.0, cy = sy / 2.0;
    const double cj = (j > cy) ? j - sy : j;
    const double ci = (i > cx) ? i - sx : i;
    double weight = 0;
    for (int w = 0; w < n; ++w) {
        const double zj = (ci * sint[w] * cosp[w] + cj * sint[w] * sinp[w]) / sx;
        double value = diameter * M_PI * zj;
        if (value == 0) value = 1;
        else value = sin(value) / value;
        weight += value;
    }
    if (weight < 0.005) weight = (double)0.005f;
    weight *= diameter;
    mask[(size_t)j * sx + i] = (float)(1 / weight);
}

// Ramp x Hamming along x (the axis perpendicular to a vertical tilt axis):
// H(k) = (pi / nproj) * |k| / sx * window(|k| / sx), window = 0.54 + 0.46 cos(pi f / fc)
// for f = |k|/sx <= fc = cutoff/2, 0 above; hamming == 0 keeps the bare ramp up to fc.
__global__ void k_ramp(int sx, int sy, int hamming, double cutoff, double ang_step, float* __restrict__ mask)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= sx) return;
    const double ci = (i > sx / 2.0) ? i - sx : i;
    const double f = fabs(ci) / sx;
    const double fc = 0.5 * cutoff;
    double h = 0.0;
    if (f <= fc) {
        const double win = hamming ? 0.54 + 0.46 * cos(M_PI * f / fc) : 1.0;
        h = ang_step * f * win;
    }
    mask[(size_t)j * sx + i] = (float)h;
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

template <int MODE>
static hipError_t row_transform(hipStream_t st, const void* in, const int* idx, void* out, const float* mask,
                                const float2* tw, int n, int nrows, int rows_per_img, size_t img_px, int inverse,
                                float scale)
{
    if (pow2(n) && n <= 4096) {
        const size_t lds = sizeof(float2) * 2 * n;
        k_fft_row<MODE><<<nrows, 256, lds, st>>>(in, idx, out, mask, tw, n, ilog2(n), rows_per_img, img_px, inverse,
                                                 scale);
    } else {
        k_dft_row<MODE><<<dim3((n + 255) / 256, nrows), 256, 0, st>>>(in, idx, out, mask, tw, n, rows_per_img, img_px,
                                                                     inverse, scale);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fourier shell correlation, ResolutionEstimation.fsc(even, odd)
// (tomography/ResolutionEstimation.java:456-557).  Each volume is multiplied by
// a soft spherical mask, transformed in 3D (FP32: x rows, a per-plane
// transpose, z rows, a full transpose, y rows; the result is laid out
// [x][z][y]), and the shell sums are accumulated in FP64.
// ---------------------------------------------------------------------------

// H = data * mask (:470-497): mask 1 within radius D (the slice count) of the
// centre (W/2, H/2, D/2), a raised-cosine fall-off over the next 20 voxels, 0
// beyond; dist^2 summed as dx^2 + (dz^2 + dy^2).  Volume in the device layout
// [y][z][x].
__global__ __launch_bounds__(256) void k_fsc_load(const float* __restrict__ vol, float2* __restrict__ out, int W,
                                                  int H, int D)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y, y = blockIdx.z;
    if (x >= W) return;
    double dz = z - D / 2.0;
    dz *= dz;
    double dy = y - H / 2.0;
    dy *= dy;
    double dist = x - W / 2.0;
    dist *= dist;
    dist += dz + dy;
    dist = sqrt(dist);
    float m = 0.0f;
    if (dist < D) m = 1.0f;
    else if (dist < D + 20) m = (float)(1 + cos((dist - D) * M_PI / 20)) * .5f;
    const size_t k = ((size_t)y * D + z) * W + x;
    out[k] = make_float2(vol[k] * m, 0.0f);
}

// convertToFrequency (:699-701)
__device__ inline double fsc_freq(int i, int n) { return (double)((i <= (n >> 1)) ? i : i - n) / n; }

// Shell sums over the transforms E and O ([x][z][y]): per shell, the real part
// of E conj(O) (formed in float as multConjSecond does), |E|^2 and |O|^2 (the
// float sums of squares, square-rooted and squared again in double), |E - O|
// and the count.  Block partials in LDS, then FP64 atomics into sums[5][R].
__global__ __launch_bounds__(256) void k_fsc_shells(const float2* __restrict__ E, const float2* __restrict__ O,
                                                    int W, int H, int D, int R, double* __restrict__ sums)
{
    extern __shared__ double sh[];  // [5][R]
    for (int k = threadIdx.x; k < 5 * R; k += blockDim.x) sh[k] = 0.0;
    __syncthreads();
    const size_t n = (size_t)W * H * D;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(k % H);
        const size_t r = k / H;
        const int z = (int)(r % D), x = (int)(r / D);
        double tz = fsc_freq(z, D);
        tz *= tz;
        double ty = fsc_freq(y, H);
        ty *= ty;
        double tx = fsc_freq(x, W);
        tx *= tx;
        const double w = sqrt(tx + ty + tz);
        const long wi = (long)floor(w * W + 0.5);  // Math.round
        if (wi >= R) continue;
        const float2 e = E[k], o = O[k];
        const double ae = sqrt((double)(e.x * e.x + e.y * e.y));
        const double ao = sqrt((double)(o.x * o.x + o.y * o.y));
        const float re = e.x * o.x + e.y * o.y;
        const double d0 = (double)(e.x - o.x), d1 = (double)(e.y - o.y);
        atomicAdd(&sh[wi], (double)re);
        atomicAdd(&sh[R + wi], ae * ae);
        atomicAdd(&sh[2 * R + wi], ao * ao);
        atomicAdd(&sh[3 * R + wi], sqrt(d0 * d0 + d1 * d1));
        atomicAdd(&sh[4 * R + wi], 1.0);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 5 * R; k += blockDim.x)
        if (sh[k] != 0.0) atomicAdd(&sums[k], sh[k]);
}

// Row transforms of many rows: the DFT path's grid.y is bounded, so rows go in
// chunks there.
static hipError_t rows_fft(hipStream_t st, const float2* in, float2* out, const float2* tw, int n, size_t nrows)
{
    if (pow2(n) && n <= 4096) {
        const size_t lds = sizeof(float2) * 2 * n;
        for (size_t r0 = 0; r0 < nrows; r0 += 0x7fffffff) {
            const size_t m = std::min<size_t>(nrows - r0, 0x7fffffff);
            k_fft_row<kComplex><<<(unsigned)m, 256, lds, st>>>(in + r0 * n, nullptr, out + r0 * n, nullptr, tw, n,
                                                               ilog2(n), 1, 0, 0, 1.0f);
        }
    } else {
        for (size_t r0 = 0; r0 < nrows; r0 += 65535) {
            const size_t m = std::min<size_t>(nrows - r0, 65535);
            k_dft_row<kComplex><<<dim3((n + 255) / 256, (unsigned)m), 256, 0, st>>>(in + r0 * n, nullptr,
                                                                                   out + r0 * n, nullptr, tw, n, 1,
                                                                                   0, 0, 1.0f);
        }
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SSNR per-tilt vectors, ResolutionEstimation.addToVectors1D
// (ResolutionEstimation.java:581-697, the computeVSSNR = false path): for
// each tilt, NaN -> 0 in the measured and re-projected images of the signal
// and noise reconstructions, the differences measured - projected (float),
// the 2-D power spectra of difference and projection for both (FP32
// transforms, powers in FP64), and ring sums: a frequency at radius w*W adds
// to rings [max(0, ceil(w*W - ringWidth)), floor(w*W)).
// ---------------------------------------------------------------------------

// out[4b + {0,1,2,3}] = signal - thSignal, thSignal, noise - thNoise, thNoise
__global__ __launch_bounds__(256) void k_ssnr_prep(const float* __restrict__ sexp, const float* __restrict__ sth,
                                                   const float* __restrict__ nexp, const float* __restrict__ nth,
                                                   size_t npx, int t0, float2* __restrict__ out)
{
    const int b = blockIdx.y;
    const size_t t = (size_t)(t0 + b);
    auto nz = [](float v) { return v != v ? 0.0f : v; };  // convertNaNToZeros
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < npx; k += (size_t)gridDim.x * blockDim.x) {
        const float s = nz(sexp[t * npx + k]), ts = nz(sth[t * npx + k]);
        const float n = nz(nexp[t * npx + k]), tn = nz(nth[t * npx + k]);
        out[(4 * (size_t)b + 0) * npx + k] = make_float2(s - ts, 0.0f);
        out[(4 * (size_t)b + 1) * npx + k] = make_float2(ts, 0.0f);
        out[(4 * (size_t)b + 2) * npx + k] = make_float2(n - tn, 0.0f);
        out[(4 * (size_t)b + 3) * npx + k] = make_float2(tn, 0.0f);
    }
}

// spectra: nb quads of images laid out [x][y] (ph fastest); sums[5][W]:
// S_S (thSignal), S_N (signal - thSignal), N_S (thNoise), N_N (noise -
// thNoise), K (count).
__global__ __launch_bounds__(256) void k_ssnr_rings(const float2* __restrict__ F, int pw, int ph, int nb,
                                                    int ring_width, double* __restrict__ sums)
{
    extern __shared__ double sh[];  // [5][pw]
    for (int k = threadIdx.x; k < 5 * pw; k += blockDim.x) sh[k] = 0.0;
    __syncthreads();
    const size_t npx = (size_t)pw * ph;
    const size_t total = npx * nb;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t b = e / npx, r = e - b * npx;
        const int x = (int)(r / ph), y = (int)(r - (size_t)x * ph);
        double ty = fsc_freq(y, ph);
        ty *= ty;
        double tx = fsc_freq(x, pw);
        tx *= tx;
        const double wi = sqrt(tx + ty) * pw;
        const int l0 = max(0, (int)ceil(wi - ring_width));
        const int lf = (int)floor(wi);
        if (l0 >= lf) continue;
        const float2* q = F + 4 * b * npx + r;
        auto pw2 = [](float2 v) { return (double)v.x * v.x + (double)v.y * v.y; };
        const double sn = pw2(q[0]), ss = pw2(q[npx]), nn = pw2(q[2 * npx]), ns = pw2(q[3 * npx]);
        for (int l = l0; l < lf; ++l) {
            atomicAdd(&sh[l], ss);
            atomicAdd(&sh[pw + l], sn);
            atomicAdd(&sh[2 * pw + l], ns);
            atomicAdd(&sh[3 * pw + l], nn);
            atomicAdd(&sh[4 * pw + l], 1.0);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 5 * pw; k += blockDim.x)
        if (sh[k] != 0.0) atomicAdd(&sums[k], sh[k]);
}

// The per-pixel SSNR map of the VSSNR path (:648-663): ISSNR = P(thSignal) /
// P(signal - thSignal), alpha the same for the noise, SSNR = max(ISSNR/alpha -
// 1, 0) (Java's max keeps NaN), value (float)(10 log10(SSNR + 1)), placed by
// swapQuadrant2D (:720-733) with centres (W/2, H/2): the first W/2 entries of a
// row move right by W/2, the next W/2 to the start, the row to (y + H/2) mod H;
// an odd width's last column is not copied (the caller zeroes the maps).
__global__ __launch_bounds__(256) void k_vssnr_map(const float2* __restrict__ F, int pw, int ph, int nb, int t0,
                                                   float* __restrict__ maps)
{
    const size_t npx = (size_t)pw * ph;
    const size_t total = npx * nb;
    const int cx = pw / 2, cy = ph / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t b = e / npx, r = e - b * npx;
        const int x = (int)(r / ph), y = (int)(r - (size_t)x * ph);
        if (x >= 2 * cx) continue;
        const float2* q = F + 4 * b * npx + r;
        auto pw2 = [](float2 v) { return (double)v.x * v.x + (double)v.y * v.y; };
        const double issnr = pw2(q[npx]) / pw2(q[0]);
        const double alpha = pw2(q[3 * npx]) / pw2(q[2 * npx]);
        const double aux = issnr / alpha - 1.0;
        const double ssnr = (aux != aux) ? aux : (aux > 0.0 ? aux : 0.0);
        int yy = y + cy;
        if (yy >= ph) yy -= ph;
        const int xx = (x < cx) ? x + cx : x - cx;
        maps[((size_t)(t0 + b) * ph + yy) * pw + xx] = (float)(10.0 * log10(ssnr + 1.0));
    }
}

}  // namespace tjf

namespace tjk {
using namespace tjf;

constexpr int kSsnrBatch = 8;  // tilts per pass

size_t ssnr_workspace(int pw, int ph)
{
    return sizeof(float2) * (2 * 4 * (size_t)kSsnrBatch * pw * ph + (size_t)pw + ph) + 256;
}

// Ring sums of all n tilts into sums[5][pw] (zeroed here).  sexp/nexp: the
// measured images, sth/nth: the projections of the reconstructions, n x ph x pw.
hipError_t ssnr(hipStream_t st, const float* sexp, const float* sth, const float* nexp, const float* nth, int pw,
                int ph, int n, int ring_width, void* ws, size_t ws_size, double* sums, float* maps)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (ws_size < ssnr_workspace(pw, ph)) return hipErrorInvalidValue;
    if (sizeof(double) * 5 * pw > 64 * 1024) return hipErrorInvalidValue;
    const size_t npx = (size_t)pw * ph;
    float2* A = (float2*)ws;
    float2* B = A + 4 * (size_t)kSsnrBatch * npx;
    float2* tww = B + 4 * (size_t)kSsnrBatch * npx;
    float2* twh = tww + pw;
    k_twiddles<<<(pw + 255) / 256, 256, 0, st>>>(pw, tww);
    k_twiddles<<<(ph + 255) / 256, 256, 0, st>>>(ph, twh);
    hipError_t e = hipMemsetAsync(sums, 0, sizeof(double) * 5 * pw, st);
    if (e == hipSuccess && maps) e = hipMemsetAsync(maps, 0, sizeof(float) * npx * n, st);
    if (e != hipSuccess) return e;
    for (int t0 = 0; t0 < n; t0 += kSsnrBatch) {
        const int m = std::min(kSsnrBatch, n - t0);
        const unsigned bx = (unsigned)std::min<size_t>((npx + 255) / 256, 1024);
        k_ssnr_prep<<<dim3(bx, m), 256, 0, st>>>(sexp, sth, nexp, nth, npx, t0, A);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // rows along x, transpose each image to [x][y], rows along y
        if ((e = rows_fft(st, A, B, tww, pw, (size_t)4 * m * ph)) != hipSuccess) return e;
        k_transpose<<<dim3((pw + 31) / 32, (ph + 31) / 32, 4 * m), dim3(32, 8), 0, st>>>(B, A, ph, pw);
        if ((e = rows_fft(st, A, B, twh, ph, (size_t)4 * m * pw)) != hipSuccess) return e;
        const unsigned blocks = (unsigned)std::min<size_t>((npx * m + 255) / 256, 1024);
        k_ssnr_rings<<<blocks, 256, sizeof(double) * 5 * pw, st>>>(B, pw, ph, m, ring_width, sums);
        if (maps) k_vssnr_map<<<blocks, 256, 0, st>>>(B, pw, ph, m, t0, maps);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipGetLastError();
}

// Workspace of tjk::fsc: three complex volumes and the twiddle tables.
size_t fsc_workspace(int W, int H, int D)
{
    const size_t n = (size_t)W * H * D;
    return sizeof(float2) * (3 * n + (size_t)W + H + D) + 256;
}

// ResolutionEstimation.fsc (:456-557) of two device volumes ([y][z][x]) into
// sums[5][R] (R = W/2 + 1; zeroed here): the caller forms the curves.
hipError_t fsc(hipStream_t st, const float* even, const float* odd, int W, int H, int D, void* ws, size_t ws_size,
               double* sums)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (ws_size < fsc_workspace(W, H, D)) return hipErrorInvalidValue;
    if (H > 65535 || D > 65535 || sizeof(double) * 5 * (W / 2 + 1) > 64 * 1024) return hipErrorInvalidValue;
    const size_t n = (size_t)W * H * D;
    const int R = W / 2 + 1;
    float2* A = (float2*)ws;
    float2* B = A + n;
    float2* X = B + n;
    float2* tww = X + n;
    float2* twh = tww + W;
    float2* twd = twh + H;
    k_twiddles<<<(W + 255) / 256, 256, 0, st>>>(W, tww);
    k_twiddles<<<(H + 255) / 256, 256, 0, st>>>(H, twh);
    k_twiddles<<<(D + 255) / 256, 256, 0, st>>>(D, twd);
    hipError_t e = hipMemsetAsync(sums, 0, sizeof(double) * 5 * R, st);
    if (e != hipSuccess) return e;
    // per volume: mask -> A; x rows A -> B; planes [z][x] -> [x][z]: B -> A;
    // z rows A -> B; [y][(x,z)] -> [(x,z)][y]: B -> A; y rows A -> result
    // (even: X, odd: B, free again by then).
    for (int v = 0; v < 2; ++v) {
        float2* res = (v == 0) ? X : B;
        k_fsc_load<<<dim3((W + 255) / 256, D, H), 256, 0, st>>>(v == 0 ? even : odd, A, W, H, D);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = rows_fft(st, A, B, tww, W, (size_t)H * D)) != hipSuccess) return e;
        k_transpose<<<dim3((W + 31) / 32, (D + 31) / 32, H), dim3(32, 8), 0, st>>>(B, A, D, W);
        if ((e = rows_fft(st, A, B, twd, D, (size_t)H * W)) != hipSuccess) return e;
        k_transpose<<<dim3((unsigned)(((size_t)W * D + 31) / 32), (H + 31) / 32, 1), dim3(32, 8), 0, st>>>(
            B, A, H, W * D);
        if ((e = rows_fft(st, A, res, twh, H, (size_t)W * D)) != hipSuccess) return e;
    }
    const size_t lds = sizeof(double) * 5 * R;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
    k_fsc_shells<<<blocks, 256, lds, st>>>(X, B, W, H, D, R, sums);
    return hipGetLastError();
}

// workspace: two complex buffers of kMaxBatch images + twiddle tables (pw + ph entries)
size_t weighting_workspace(int pw, int ph, int n)
{
    const size_t nb = (size_t)std::min(n, kMaxBatch);
    return sizeof(float2) * (2 * nb * (size_t)pw * ph + (size_t)pw + (size_t)ph) + 256;
}

hipError_t weighting_init(hipStream_t st, int sx, int sy, double diameter, const double* tilt_deg,
                          const double* axis_deg, int n, float* mask)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    // FFTWeighting.initWeighting (:67-81): sin(tilt), cos(axis), sin(axis) per
    // image; Java 8's Math.toRadians is angdeg / 180.0 * PI.
    std::vector<double> trig(3 * (size_t)n);
    for (int w = 0; w < n; ++w) {
        const double ax = axis_deg ? axis_deg[w] : 0.0;
        trig[w] = std::sin(tilt_deg[w] / 180.0 * M_PI);
        trig[n + w] = std::cos(ax / 180.0 * M_PI);
        trig[2 * n + w] = std::sin(ax / 180.0 * M_PI);
    }
    double* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, sizeof(double) * 3 * n);
    if (e != hipSuccess) return e;
    e = hipMemcpy(d, trig.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_radermacher<<<dim3((sx + 255) / 256, sy), 256, 0, st>>>(sx, sy, diameter, d, d + n, d + 2 * n, n, mask);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    return e;
}

hipError_t ramp_init(hipStream_t st, int sx, int sy, int hamming, double cutoff, double ang_step, float* mask)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    k_ramp<<<dim3((sx + 255) / 256, sy), 256, 0, st>>>(sx, sy, hamming, cutoff, ang_step, mask);
    return hipGetLastError();
}

// Weights n images (exps[idx[s]]) into out[s] (float, [n][ph][pw]).
hipError_t weighting_apply(hipStream_t st, int pw, int ph, int n, const float* exps, const int* idx, const float* mask,
                           float* out, void* ws, size_t ws_size)
{
    (void)hipGetLastError();  // drop a stale error left by an earlier runtime call
    if (n <= 0) return hipSuccess;
    if (ws_size < weighting_workspace(pw, ph, n)) return hipErrorInvalidValue;
    const size_t npx = (size_t)pw * ph;
    const int nb = std::min(n, kMaxBatch);
    float2* A = (float2*)ws;
    float2* B = A + nb * npx;
    float2* tww = B + nb * npx;
    float2* twh = tww + pw;
    k_twiddles<<<(pw + 255) / 256, 256, 0, st>>>(pw, tww);
    k_twiddles<<<(ph + 255) / 256, 256, 0, st>>>(ph, twh);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const float scale = (float)(1.0 / (double)npx);
    for (int b0 = 0; b0 < n; b0 += nb) {
        const int m = std::min(nb, n - b0);
        // rows of the images: forward transform along x
        e = row_transform<kRealIn>(st, exps, idx + b0, A, nullptr, tww, pw, m * ph, ph, npx, 0, 1.0f);
        if (e != hipSuccess) return e;
        k_transpose<<<dim3((pw + 31) / 32, (ph + 31) / 32, m), dim3(32, 8), 0, st>>>(A, B, ph, pw);
        // columns: forward along y, mask, inverse along y
        e = row_transform<kComplexMask>(st, B, nullptr, A, mask, twh, ph, m * pw, pw, npx, 0, 1.0f);
        if (e != hipSuccess) return e;
        e = row_transform<kComplex>(st, A, nullptr, B, nullptr, twh, ph, m * pw, pw, npx, 1, 1.0f);
        if (e != hipSuccess) return e;
        k_transpose<<<dim3((ph + 31) / 32, (pw + 31) / 32, m), dim3(32, 8), 0, st>>>(B, A, pw, ph);
        // inverse along x, real part, 1/(W*H)
        e = row_transform<kRealOut>(st, A, nullptr, out + (size_t)b0 * npx, nullptr, tww, pw, m * ph, ph, npx, 1,
                                    scale);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace tjk
