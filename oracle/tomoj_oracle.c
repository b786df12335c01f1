/*
 * tomoj_oracle.c -- CPU restatement of TomoJ's parallel-beam projector path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * path in tomoj_amd/csrc.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or as the timed CPU
 * baseline).  Nothing in the product links against it.
 *
 * Two restatements live here:
 *   - "faithful" (FP32): mirrors TomoJ's Java arithmetic operation by operation
 *     (float accumulators, double positions, float casts at the same points,
 *     no fused multiply-add), so that a correct GPU path can be bit-identical.
 *   - "f64": the same algorithm with every value held in double; it is the
 *     FP64 host oracle the error of each result is stated against.
 *
 * Reference files followed (paths relative to src/TomoJ_Base/src/main/java/fr/curie/tomoj/):
 *   tomography/projectors/VoxelProjector3D.java   projectJosephPartial :671-821,
 *        project :333-376, projectAndDifference :387-401, difference :414-433,
 *        backproject :155-208, backProjectPartial :221-308
 *   tomography/TomoReconstruction2.java           getPixel :154-190, WBP :499-538,
 *        OSSART :548-634, osartIteration :636-733, orderShuffleART :735-781
 *   tomography/filters/FFTWeighting.java          createWeightingMask :31-65, weighting :87-112
 *   src/Utility/.../fr/curie/utils/MatrixUtils.java eulerAngles2Matrix :226-251
 *   tomography/TiltSeries.java                    applyThreshold :720-806 (projection supply)
 *   TomoJ_AdvancedReconstruction/.../projectors/BgARTVoxelProjector3D.java collapseFilter :34-78,
 *        medianFilter3Dcross :123-151
 *   src/Filters/.../fr/curie/filters/TotalVariationDenoising_3D.java denoise3D :84-202,
 *        denoise2D :206-260 (the step TVMProjector3D.endOfIteration runs)
 *   third party: ImageJ 1.53a ImageProcessor.getInterpolatedValue (not in the
 *   reference tree; restated from its published source, see ij_interpolated_value).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  FMA contraction
 * must stay off: Java never fuses a*b+c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

#define OR_ACCURACY ((double)1e-10f) /* VoxelProjector3D.java:33, a float widened to double */
#define JAVA_FLOAT_MAX 3.4028234663852886e38

typedef struct {
    int vw, vh, vd;       /* volume width (x), height (y), depth (#slices, z)      */
    int pw, ph;           /* detector width, height                                 */
    double pc[4];         /* Projector.projectionCenters[0..3]   (Projector.java:52-58) */
    double vc[3];         /* Projector.volumeCenters[0..2]       (Projector.java:60-65) */
    double scale;         /* VoxelProjector3D.scale              (VoxelProjector3D.java:89-91) */
    int positivity;       /* positivityConstraint                (VoxelProjector3D.java:102-104) */
    int loc;              /* longObjectCompensation              (VoxelProjector3D.java:78-80) */
    int nthreads;         /* worker threads (TomoJ: Prefs.getThreads()) ; <=0 -> 1 */
} or_geom;

/* ------------------------------------------------------------------------- */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------- */

/* Java 8 Math.toRadians: angdeg / 180.0 * PI */
static double java8_to_radians(double deg) { return deg / 180.0 * M_PI; }

/* MatrixUtils.eulerAngles2Matrix (MatrixUtils.java:226-251); E row-major 3x3. */
void or_euler_matrix(double phi, double theta, double psi, double *E)
{
    phi = java8_to_radians(phi);
    theta = java8_to_radians(theta);
    psi = java8_to_radians(psi);
    double cphi = cos(phi), sphi = sin(phi);
    double ctheta = cos(theta), stheta = sin(theta);
    double cpsi = cos(psi), spsi = sin(psi);
    E[0] = cpsi * ctheta * cphi - spsi * sphi;
    E[1] = cpsi * ctheta * sphi + spsi * cphi;
    E[2] = -cpsi * stheta;
    E[3] = -spsi * ctheta * cphi - cpsi * sphi;
    E[4] = -spsi * ctheta * sphi + cpsi * cphi;
    E[5] = spsi * stheta;
    E[6] = stheta * cphi;
    E[7] = stheta * sphi;
    E[8] = ctheta;
}

/* AffineAlignment.getEulerMatrix with tiltAxisVertical (AffineAlignment.java:382-395):
 * eulerAngles2Matrix(ta, tilt, -ta), ta = 0 when the axis is vertical. */
void or_tilt_euler(double tilt_deg, double tilt_axis_deg, int axis_vertical, double *E)
{
    double ta = axis_vertical ? 0.0 : tilt_axis_deg;
    or_euler_matrix(ta, tilt_deg, -ta, E);
}

/* ------------------------------------------------------------------------- */
/* Samplers                                                                  */
/* ------------------------------------------------------------------------- */

/* TomoReconstruction2.getPixel(double x, double y, int z) (TomoReconstruction2.java:154-190).
 * vol is [z][y][x].  The callers' inclusion test keeps x < W and y < H, so
 * every index stays inside the slice; the guards below only protect direct
 * calls with out-of-range coordinates. */
static inline float getpixel_f32(const float *vol, int W, int H, int D, double x, double y, int z)
{
    const int xbase = (x < 0) ? 0 : (int)x;
    const int ybase = (y < 0) ? 0 : (int)y;
    const float *img = vol + (size_t)z * W * H;
    const long slice = (long)W * H;
    if (xbase == W - 1 || ybase == H - 1 || z == D - 1) {
        long idx = xbase + (long)ybase * W;
        return (idx < slice) ? img[idx] : 0.0f;
    }
    const float xf = (x < 0) ? 0.0f : (float)x - (float)xbase;
    const float yf = (y < 0) ? 0.0f : (float)y - (float)ybase;
    const long index = (long)ybase * W + xbase;
    if (index + W + 1 >= slice) return 0.0f;
    const float xyz = img[index];
    const float xy1z = img[index + W];
    const float lower = xyz + xf * (img[index + 1] - xyz);
    return lower + yf * ((xy1z + xf * (img[index + W + 1] - xy1z)) - lower);
}

static inline double getpixel_f64(const double *vol, int W, int H, int D, double x, double y, int z)
{
    const int xbase = (x < 0) ? 0 : (int)x;
    const int ybase = (y < 0) ? 0 : (int)y;
    const double *img = vol + (size_t)z * W * H;
    const long slice = (long)W * H;
    if (xbase == W - 1 || ybase == H - 1 || z == D - 1) {
        long idx = xbase + (long)ybase * W;
        return (idx < slice) ? img[idx] : 0.0;
    }
    const double xf = (x < 0) ? 0.0 : x - xbase;
    const double yf = (y < 0) ? 0.0 : y - ybase;
    const long index = (long)ybase * W + xbase;
    if (index + W + 1 >= slice) return 0.0;
    const double xyz = img[index];
    const double xy1z = img[index + W];
    const double lower = xyz + xf * (img[index + 1] - xyz);
    return lower + yf * ((xy1z + xf * (img[index + W + 1] - xy1z)) - lower);
}

/* ImageJ 1.53a ImageProcessor.getInterpolatedValue(double x, double y) on a
 * FloatProcessor: 0 outside [-1,W)x[-1,H); bilinear in double inside, with the
 * neighbour index clamped to the last column/row in the edge band
 * (getInterpolatedEdgeValue / getEdgeValue).  Restated from ImageJ's published
 * source; the library is not vendored in the reference.  In the band
 * [-1,0) the edge path floors the coordinate and clamps both neighbours. */
static inline double ij_interpolated_value(const float *img, int W, int H, double x, double y)
{
    if (x < -1 || x >= W || y < -1 || y >= H) return 0.0;
    int xbase, ybase;
    if (x < 0.0 || y < 0.0 || x >= W - 1.0 || y >= H - 1.0) {
        xbase = (int)floor(x);
        ybase = (int)floor(y);
    } else {
        xbase = (int)x;
        ybase = (int)y;
    }
    double xf = x - xbase, yf = y - ybase;
    if (xf < 0.0) xf = 0.0;
    if (yf < 0.0) yf = 0.0;
    int x0 = xbase <= 0 ? 0 : (xbase >= W ? W - 1 : xbase);
    int x1 = xbase + 1 <= 0 ? 0 : (xbase + 1 >= W ? W - 1 : xbase + 1);
    int y0 = ybase <= 0 ? 0 : (ybase >= H ? H - 1 : ybase);
    int y1 = ybase + 1 <= 0 ? 0 : (ybase + 1 >= H ? H - 1 : ybase + 1);
    double ll = img[(long)y0 * W + x0];
    double lr = img[(long)y0 * W + x1];
    double ur = img[(long)y1 * W + x1];
    double ul = img[(long)y1 * W + x0];
    double upper = ul + xf * (ur - ul);
    double lower = ll + xf * (lr - ll);
    return lower + yf * (upper - lower);
}

static inline double ij_interpolated_value_f64(const double *img, int W, int H, double x, double y)
{
    if (x < -1 || x >= W || y < -1 || y >= H) return 0.0;
    int xbase, ybase;
    if (x < 0.0 || y < 0.0 || x >= W - 1.0 || y >= H - 1.0) {
        xbase = (int)floor(x);
        ybase = (int)floor(y);
    } else {
        xbase = (int)x;
        ybase = (int)y;
    }
    double xf = x - xbase, yf = y - ybase;
    if (xf < 0.0) xf = 0.0;
    if (yf < 0.0) yf = 0.0;
    int x0 = xbase <= 0 ? 0 : (xbase >= W ? W - 1 : xbase);
    int x1 = xbase + 1 <= 0 ? 0 : (xbase + 1 >= W ? W - 1 : xbase + 1);
    int y0 = ybase <= 0 ? 0 : (ybase >= H ? H - 1 : ybase);
    int y1 = ybase + 1 <= 0 ? 0 : (ybase + 1 >= H ? H - 1 : ybase + 1);
    double ll = img[(long)y0 * W + x0];
    double lr = img[(long)y0 * W + x1];
    double ur = img[(long)y1 * W + x1];
    double ul = img[(long)y1 * W + x0];
    double upper = ul + xf * (ur - ul);
    double lower = ll + xf * (lr - ll);
    return lower + yf * (upper - lower);
}

double or_ij_interpolated_value(const float *img, int W, int H, double x, double y)
{
    return ij_interpolated_value(img, W, H, x, y);
}

float or_getpixel(const float *vol, int W, int H, int D, double x, double y, int z)
{
    return getpixel_f32(vol, W, H, D, x, y, z);
}

/* ------------------------------------------------------------------------- */
/* Forward projection (Joseph, slice stepping)                               */
/* ------------------------------------------------------------------------- */

typedef struct {
    double icf, zcf, divy, yfx, zfx, divx, dirx, diry, norminc;
} fp_factors;

/* Per-projection constants of projectJosephPartial (VoxelProjector3D.java:682-694). */
static fp_factors fp_make_factors(const double *r)
{
    fp_factors f;
    f.icf = r[3] / r[0];
    f.zcf = (r[5] - (r[3] * r[2] / r[0]));
    f.divy = r[4] - (r[3] * r[1] / r[0]);
    f.yfx = r[1];
    f.zfx = r[2];
    f.divx = r[0];
    f.diry = -(r[5] - r[3] * r[2] / r[0]) / (r[4] - r[3] * r[1] / r[0]);
    f.dirx = (r[1] * (-f.diry) - r[2]) / r[0];
    f.norminc = sqrt(f.dirx * f.dirx + f.diry * f.diry + 1);
    return f;
}

void or_fp_factors(const double *E, double *out9)
{
    fp_factors f = fp_make_factors(E);
    out9[0] = f.icf; out9[1] = f.zcf; out9[2] = f.divy; out9[3] = f.yfx; out9[4] = f.zfx;
    out9[5] = f.divx; out9[6] = f.dirx; out9[7] = f.diry; out9[8] = f.norminc;
}

/* projectJosephPartial for rows [starty,endy) of one projection
 * (VoxelProjector3D.java:671-821).  Returns the task's norm_max. */
static float fp_task_f32(const or_geom *g, const float *vol, const double *E,
                         float *proj, float *norm, int starty, int endy)
{
    const fp_factors f = fp_make_factors(E);
    const double vcx = g->vc[0], vcy = g->vc[1], vcz = g->vc[2];
    /* volumeWidth + ACCURACY is int + float, i.e. float arithmetic (:700-701):
     * the borders are exactly W and H. */
    const double bxmax = (double)((float)g->vw + 1e-10f), bymax = (double)((float)g->vh + 1e-10f);
    const int psx = g->pw;
    float norm_max = 0;
    long jj = (long)(starty - 1) * psx;
    double jcentered = (starty - 1.0 - g->pc[1]) / g->scale;
    const double jinc = 1.0 / g->scale;
    for (int j = starty; j < endy; j++) {
        jj += psx;
        jcentered += jinc;
        if (j >= 0 && j < g->ph) {
            for (int i = 0; i < g->pw; i++) {
                double icentered = (i - g->pc[0]) / g->scale;
                float ray_sum = 0, norm_sum = 0;
                double ijpart = jcentered - f.icf * icentered;
                double oldy = (ijpart - f.zcf * (-1 - vcz)) / f.divy;
                double oldx = icentered - f.yfx * oldy - f.zfx * (-1 - vcz);
                oldx /= f.divx;
                int entered = 0;
                for (int z = 0; z < g->vd; z++) {
                    double xc = oldx + f.dirx;
                    double yc = oldy + f.diry;
                    double x = xc + vcx;
                    double y = yc + vcy;
                    if (x + OR_ACCURACY >= 0 && x < bxmax && y + OR_ACCURACY >= 0 && y < bymax) {
                        float value = getpixel_f32(vol, g->vw, g->vh, g->vd, x, y, z);
                        ray_sum += value;
                        norm_sum = (float)((double)norm_sum + f.norminc);
                        entered = 1;
                    } else if (entered) {
                        norm_sum = (float)((double)norm_sum + f.norminc);
                        break;
                    }
                    oldx = xc;
                    oldy = yc;
                }
                if (norm_sum > 0) {
                    proj[jj + i] = (float)((double)ray_sum * f.norminc);
                    norm[jj + i] = norm_sum;
                    if (norm_sum > norm_max) norm_max = norm_sum;
                } else {
                    proj[jj + i] = NAN;
                    norm[jj + i] = 0;
                }
            }
        }
    }
    return norm_max;
}

static double fp_task_f64(const or_geom *g, const double *vol, const double *E,
                          double *proj, double *norm, int starty, int endy)
{
    const fp_factors f = fp_make_factors(E);
    const double vcx = g->vc[0], vcy = g->vc[1], vcz = g->vc[2];
    /* volumeWidth + ACCURACY is int + float, i.e. float arithmetic (:700-701):
     * the borders are exactly W and H. */
    const double bxmax = (double)((float)g->vw + 1e-10f), bymax = (double)((float)g->vh + 1e-10f);
    const int psx = g->pw;
    double norm_max = 0;
    long jj = (long)(starty - 1) * psx;
    double jcentered = (starty - 1.0 - g->pc[1]) / g->scale;
    const double jinc = 1.0 / g->scale;
    for (int j = starty; j < endy; j++) {
        jj += psx;
        jcentered += jinc;
        if (j >= 0 && j < g->ph) {
            for (int i = 0; i < g->pw; i++) {
                double icentered = (i - g->pc[0]) / g->scale;
                double ray_sum = 0, norm_sum = 0;
                double ijpart = jcentered - f.icf * icentered;
                double oldy = (ijpart - f.zcf * (-1 - vcz)) / f.divy;
                double oldx = icentered - f.yfx * oldy - f.zfx * (-1 - vcz);
                oldx /= f.divx;
                int entered = 0;
                for (int z = 0; z < g->vd; z++) {
                    double xc = oldx + f.dirx;
                    double yc = oldy + f.diry;
                    double x = xc + vcx;
                    double y = yc + vcy;
                    if (x + OR_ACCURACY >= 0 && x < bxmax && y + OR_ACCURACY >= 0 && y < bymax) {
                        ray_sum += getpixel_f64(vol, g->vw, g->vh, g->vd, x, y, z);
                        norm_sum += f.norminc;
                        entered = 1;
                    } else if (entered) {
                        norm_sum += f.norminc;
                        break;
                    }
                    oldx = xc;
                    oldy = yc;
                }
                if (norm_sum > 0) {
                    proj[jj + i] = ray_sum * f.norminc;
                    norm[jj + i] = norm_sum;
                    if (norm_sum > norm_max) norm_max = norm_sum;
                } else {
                    proj[jj + i] = NAN;
                    norm[jj + i] = 0;
                }
            }
        }
    }
    return norm_max;
}

/* First detector row computed by VoxelProjector3D.project (:338, :347). */
static int fp_start_row(const or_geom *g, int startY)
{
    int offset = (g->ph > g->vh && g->scale == 1) ? (g->ph - g->vh) / 2 : 0;
    return (int)floor((startY + offset) * g->scale + 0.5);
}

/* VoxelProjector3D.project for one projection: tasks of 2 detector rows from
 * round((startY+offset)*scale) to the detector height (:346-361).  proj/norm
 * must be zero-filled by the caller (fresh FloatProcessors). */
void or_project_one(const or_geom *g, const float *vol, const double *E,
                    float *proj, float *norm, int startY, int endY)
{
    (void)endY; /* TomoJ projects to the detector height whatever endY is (:348) */
    const int start = fp_start_row(g, startY), end = g->ph;
    const int ntask = (end > start) ? (end - start + 1) / 2 : 0;
    float nmax = 0;
    int nt = g->nthreads > 0 ? g->nthreads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(max : nmax)
    for (int t = 0; t < ntask; t++) {
        int sy = start + 2 * t;
        int ey = (sy + 2 < g->ph) ? sy + 2 : g->ph;
        float m = fp_task_f32(g, vol, E, proj, norm, sy, ey);
        if (m > nmax) nmax = m;
    }
    if (g->loc) {
        /* Each Java task overwrites the whole norm image with its own max
         * (:814-819); with every row inside the volume all tasks hold the same
         * max, which is the projection-wide max used here. */
        for (long k = 0; k < (long)g->pw * g->ph; k++) norm[k] = nmax;
    }
}

void or_project_one_f64(const or_geom *g, const double *vol, const double *E,
                        double *proj, double *norm, int startY, int endY)
{
    (void)endY;
    const int start = fp_start_row(g, startY), end = g->ph;
    const int ntask = (end > start) ? (end - start + 1) / 2 : 0;
    double nmax = 0;
    int nt = g->nthreads > 0 ? g->nthreads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(max : nmax)
    for (int t = 0; t < ntask; t++) {
        int sy = start + 2 * t;
        int ey = (sy + 2 < g->ph) ? sy + 2 : g->ph;
        double m = fp_task_f64(g, vol, E, proj, norm, sy, ey);
        if (m > nmax) nmax = m;
    }
    if (g->loc)
        for (long k = 0; k < (long)g->pw * g->ph; k++) norm[k] = nmax;
}

/* VoxelProjector3D.difference (:414-433).  diff must be zero-filled. */
double or_difference(const float *exp, const float *th, const float *norm, float *diff,
                     long n, double factor)
{
    double error = 0;
    for (long pos = 0; pos < n; pos++) {
        float nn = norm[pos], e = exp[pos], t = th[pos];
        if (!isnan(e) && !isnan(t)) {
            double value = (double)((e - t) / nn) * factor;
            diff[pos] = (float)value;
            error += value * value;
        }
    }
    return error;
}

double or_difference_f64(const double *exp, const double *th, const double *norm, double *diff,
                         long n, double factor)
{
    double error = 0;
    for (long pos = 0; pos < n; pos++) {
        if (!isnan(exp[pos]) && !isnan(th[pos])) {
            double value = ((exp[pos] - th[pos]) / norm[pos]) * factor;
            diff[pos] = value;
            error += value * value;
        }
    }
    return error;
}

/* ------------------------------------------------------------------------- */
/* Backprojection (voxel driven)                                             */
/* ------------------------------------------------------------------------- */

/* backProjectPartial for volume rows [starty,endy) (VoxelProjector3D.java:221-308),
 * no mask, no global deformation. */
static void bp_task_f32(const or_geom *g, float *vol, const float *proj, const double *E,
                        int starty, int endy)
{
    const double cix = g->pc[2], ciy = g->pc[3];
    for (int k = 0; k < g->vd; k++) {
        double z = k - g->vc[2];
        float *planez = vol + (size_t)k * g->vw * g->vh;
        for (int j = starty; j < endy; j++) {
            double y = j - g->vc[1];
            for (int i = 0; i < g->vw; i++) {
                double x = i - g->vc[0];
                double ztx = z * E[2], zty = z * E[5];
                double ytx = y * E[1], yty = y * E[4];
                long jj = (long)j * g->vw;
                double xx = (x * E[0] + ytx + ztx) * g->scale + cix;
                double yy = (x * E[3] + yty + zty) * g->scale + ciy;
                if (xx >= 0 && xx < g->pw && yy >= 0 && yy < g->ph) {
                    float value = (float)ij_interpolated_value(proj, g->pw, g->ph, xx, yy);
                    if (!isnan(value)) {
                        float v = planez[i + jj] + value;
                        if (isnan(v)) v = 0;
                        if (isinf(v)) v = 0;
                        if (g->positivity && v < 0) v = 0;
                        planez[i + jj] = v;
                    }
                }
            }
        }
    }
}

static void bp_task_f64(const or_geom *g, double *vol, const double *proj, const double *E,
                        int starty, int endy)
{
    const double cix = g->pc[2], ciy = g->pc[3];
    for (int k = 0; k < g->vd; k++) {
        double z = k - g->vc[2];
        double *planez = vol + (size_t)k * g->vw * g->vh;
        for (int j = starty; j < endy; j++) {
            double y = j - g->vc[1];
            for (int i = 0; i < g->vw; i++) {
                double x = i - g->vc[0];
                long jj = (long)j * g->vw;
                double xx = (x * E[0] + y * E[1] + z * E[2]) * g->scale + cix;
                double yy = (x * E[3] + y * E[4] + z * E[5]) * g->scale + ciy;
                if (xx >= 0 && xx < g->pw && yy >= 0 && yy < g->ph) {
                    double value = ij_interpolated_value_f64(proj, g->pw, g->ph, xx, yy);
                    if (!isnan(value)) {
                        double v = planez[i + jj] + value;
                        if (isnan(v) || isinf(v)) v = 0;
                        if (g->positivity && v < 0) v = 0;
                        planez[i + jj] = v;
                    }
                }
            }
        }
    }
}

/* VoxelProjector3D.backproject for one image: tasks of 2 volume rows over
 * [startY,endY), joined before the next projection (:162-206). */
void or_backproject_one(const or_geom *g, float *vol, const float *proj, const double *E,
                        int startY, int endY)
{
    const int ntask = (endY > startY) ? (endY - startY + 1) / 2 : 0;
    int nt = g->nthreads > 0 ? g->nthreads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int t = 0; t < ntask; t++) {
        int y = startY + 2 * t;
        bp_task_f32(g, vol, proj, E, y, (y + 2 < endY) ? y + 2 : endY);
    }
}

void or_backproject_one_f64(const or_geom *g, double *vol, const double *proj, const double *E,
                            int startY, int endY)
{
    const int ntask = (endY > startY) ? (endY - startY + 1) / 2 : 0;
    int nt = g->nthreads > 0 ? g->nthreads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int t = 0; t < ntask; t++) {
        int y = startY + 2 * t;
        bp_task_f64(g, vol, proj, E, y, (y + 2 < endY) ? y + 2 : endY);
    }
}

/* ------------------------------------------------------------------------- */
/* Projection order and the OS-SART loop                                     */
/* ------------------------------------------------------------------------- */

/* TomoReconstruction2.orderShuffleART (:735-781) over all images, with the
 * random first pick replaced by first_index.  eul: n x 9. */
void or_order_shuffle(const double *eul, int n, int first_index, int *order)
{
    char *todo = (char *)malloc(n);
    memset(todo, 1, n);
    int cc = 0, im = 0;
    while (cc < n) {
        if (cc == 0) {
            im = first_index;
        } else {
            double mini = DBL_MAX;
            int omini = -1;
            const double *a = eul + (size_t)im * 9 + 6;
            for (int o = 0; o < n; o++) {
                if (todo[o]) {
                    const double *b = eul + (size_t)o * 9 + 6;
                    double sum = 0;
                    sum += a[0] * b[0];
                    sum += a[1] * b[1];
                    sum += a[2] * b[2];
                    double sp = fabs(sum);
                    if (sp < mini) { mini = sp; omini = o; }
                }
            }
            im = omini;
        }
        if (todo[im]) { order[cc] = im; todo[im] = 0; cc++; }
    }
    free(todo);
}

/* orderShuffleART over a list of available images (TomoReconstruction2.java:735-781):
 * the chain runs over positions 0..nav-1 with the Euler matrix of IMAGE im
 * (getEulerMatrix(im), :741), not of the image avail[im] the position names,
 * and emits avail[im].  avail NULL: every image. */
void or_order_shuffle_avail(const double *eul, int nav, const int *avail, int first_index, int *order)
{
    if (nav <= 0) return;
    or_order_shuffle(eul, nav, first_index, order);
    if (avail)
        for (int k = 0; k < nav; k++) order[k] = avail[order[k]];
}

/* ReconstructionParameters.getAvailableIndexes (ReconstructionParameters.java:212-256)
 * for projectionType ALL (10, and any other value), EVEN (11) or ODD (12): the
 * images of that parity, at most n/2 of them.  Returns the count. */
int or_available_indexes(int n, int type, int *out)
{
    int cnt = 0, k = 0;
    char *take = (char *)malloc(n > 0 ? n : 1);
    for (int i = 0; i < n; i++) {
        if (type == 11 || type == 12) {
            const int par = (type == 11) ? (i % 2 == 0) : (i % 2 == 1);
            take[i] = (char)par;
            if (par) cnt++;
            if (cnt > n / 2) {
                take[i] = 0;
                cnt--;
            }
        } else {
            take[i] = 1;
            cnt++;
        }
    }
    for (int i = 0; i < n; i++)
        if (take[i]) out[k++] = i;
    free(take);
    return cnt;
}

/* VoxelProjector3D.convertDiffToErrorImage (VoxelProjector3D.java:315-324):
 * each pixel squared in float, NaN left as it is. */
void or_convert_diff_to_error(float *diff, long n)
{
    for (long i = 0; i < n; i++) {
        const float f = diff[i];
        if (!isnan(f)) diff[i] = f * f;
    }
}

/* One osartIteration (TomoReconstruction2.java:636-733).  projs: n x ph x pw
 * (the already-supplied projections, ts.getPixels).  err5 receives err1[0..4].
 * With a list of available images (nav of them): the order runs over the
 * list, but update (0 -> n), the flush rule "t == n - 1" and the error
 * normalisation keep the full count n, as TomoJ's do (:639, :644, :662,
 * :726), so a last partial subset is never projected, and with update 0
 * nothing is.  Returns 0, or -1 on allocation failure. */
static int osart_iteration_f32(const or_geom *g, float *vol, const float *projs, const double *eul,
                               int n, const int *avail, int nav, double lambda, int update_nb, int first_index,
                               int startY, int endY, double *err5, float *errvol)
{
    const long npx = (long)g->pw * g->ph;
    int *order = (int *)malloc(sizeof(int) * (nav > 0 ? nav : 1));
    or_order_shuffle_avail(eul, nav, avail, first_index, order);
    const int update = (update_nb == 0) ? n : update_nb;
    const float factorimg = (float)(lambda / update);
    float *th = (float *)malloc(sizeof(float) * npx * update);
    float *nm = (float *)malloc(sizeof(float) * npx * update);
    float *df = (float *)malloc(sizeof(float) * npx * update);
    double *errors = (double *)malloc(sizeof(double) * update);
    if (!order || !th || !nm || !df || !errors) return -1;
    double som = 0, errmin = JAVA_FLOAT_MAX, errmax = 0;
    int imin = 0, imax = 0, diffIndex = 0;
    for (int t = 0; t < nav; t++) {
        diffIndex++;
        if (diffIndex == update || t == n - 1) {
            const int nadd = diffIndex;
            for (int p = 0; p < nadd; p++) {
                const int idx = order[t - nadd + 1 + p];
                memset(th + p * npx, 0, sizeof(float) * npx);
                memset(nm + p * npx, 0, sizeof(float) * npx);
                memset(df + p * npx, 0, sizeof(float) * npx);
                or_project_one(g, vol, eul + (size_t)idx * 9, th + p * npx, nm + p * npx, startY, endY);
            }
            for (int p = 0; p < nadd; p++) {
                const int idx = order[t - nadd + 1 + p];
                errors[p] = or_difference(projs + (size_t)idx * npx, th + p * npx, nm + p * npx,
                                          df + p * npx, npx, (double)factorimg);
            }
            for (int e = 0; e < nadd; e++) {
                som += errors[e];
                if (errors[e] > errmax) { errmax = errors[e]; imax = order[t - nadd + e + 1]; }
                else if (errors[e] < errmin) { errmin = errors[e]; imin = order[t - nadd + e + 1]; }
            }
            /* VoxelProjector3D.backproject (:155-208): per image the
             * reconstruction's backprojection of the difference, then, with an
             * error volume, the difference squared in place and backprojected
             * into it (the reconstruction reads the unsquared difference) */
            for (int p = 0; p < nadd; p++) {
                const int idx = order[t - nadd + 1 + p];
                or_backproject_one(g, vol, df + p * npx, eul + (size_t)idx * 9, startY, endY);
                if (errvol) {
                    or_convert_diff_to_error(df + p * npx, npx);
                    or_backproject_one(g, errvol, df + p * npx, eul + (size_t)idx * 9, startY, endY);
                }
            }
            diffIndex = 0;
        }
    }
    const double wh = (double)g->vw * g->vh;
    err5[0] = sqrt(som / (double)((long)n * g->vw * g->vh));
    err5[1] = sqrt(errmin / wh);
    err5[2] = imin;
    err5[3] = sqrt(errmax / wh);
    err5[4] = imax;
    free(order); free(th); free(nm); free(df); free(errors);
    return 0;
}

static int osart_iteration_f64(const or_geom *g, double *vol, const double *projs, const double *eul,
                               int n, const int *avail, int nav, double lambda, int update_nb, int first_index,
                               int startY, int endY, double *err5)
{
    const long npx = (long)g->pw * g->ph;
    int *order = (int *)malloc(sizeof(int) * (nav > 0 ? nav : 1));
    or_order_shuffle_avail(eul, nav, avail, first_index, order);
    const int update = (update_nb == 0) ? n : update_nb;
    const double factor = lambda / update;
    double *th = (double *)malloc(sizeof(double) * npx * update);
    double *nm = (double *)malloc(sizeof(double) * npx * update);
    double *df = (double *)malloc(sizeof(double) * npx * update);
    double *errors = (double *)malloc(sizeof(double) * update);
    if (!order || !th || !nm || !df || !errors) return -1;
    double som = 0, errmin = JAVA_FLOAT_MAX, errmax = 0;
    int imin = 0, imax = 0, diffIndex = 0;
    for (int t = 0; t < nav; t++) {
        diffIndex++;
        if (diffIndex == update || t == n - 1) {
            const int nadd = diffIndex;
            for (int p = 0; p < nadd; p++) {
                const int idx = order[t - nadd + 1 + p];
                memset(th + p * npx, 0, sizeof(double) * npx);
                memset(nm + p * npx, 0, sizeof(double) * npx);
                memset(df + p * npx, 0, sizeof(double) * npx);
                or_project_one_f64(g, vol, eul + (size_t)idx * 9, th + p * npx, nm + p * npx, startY, endY);
            }
            for (int p = 0; p < nadd; p++) {
                const int idx = order[t - nadd + 1 + p];
                errors[p] = or_difference_f64(projs + (size_t)idx * npx, th + p * npx, nm + p * npx,
                                              df + p * npx, npx, factor);
            }
            for (int e = 0; e < nadd; e++) {
                som += errors[e];
                if (errors[e] > errmax) { errmax = errors[e]; imax = order[t - nadd + e + 1]; }
                else if (errors[e] < errmin) { errmin = errors[e]; imin = order[t - nadd + e + 1]; }
            }
            for (int p = 0; p < nadd; p++) {
                const int idx = order[t - nadd + 1 + p];
                or_backproject_one_f64(g, vol, df + p * npx, eul + (size_t)idx * 9, startY, endY);
            }
            diffIndex = 0;
        }
    }
    const double wh = (double)g->vw * g->vh;
    err5[0] = sqrt(som / (double)((long)n * g->vw * g->vh));
    err5[1] = sqrt(errmin / wh);
    err5[2] = imin;
    err5[3] = sqrt(errmax / wh);
    err5[4] = imax;
    free(order); free(th); free(nm); free(df); free(errors);
    return 0;
}

/* TomoReconstruction2.OSSART (:548-634): `iterations` calls of osartIteration.
 * err: iterations x 5. */
int or_ossart(const or_geom *g, float *vol, const float *projs, const double *eul, int n,
              int iterations, double lambda, int update_nb, int first_index,
              int startY, int endY, double *err)
{
    for (int it = 0; it < iterations; it++)
        if (osart_iteration_f32(g, vol, projs, eul, n, NULL, n, lambda, update_nb, first_index, startY, endY,
                                err + 5 * it, NULL))
            return -1;
    return 0;
}

/* OSSART with VoxelProjector3D's error volume (createErrorVolume, :858-864):
 * errvol (ImageStack layout like vol) accumulates the backprojected squared
 * differences of every subset of the `iterations` iterations (TomoJ saves and
 * clears it at each endOfIteration, :828-852; a caller runs one iteration per
 * call to get each iteration's volume). */
int or_ossart_errvol(const or_geom *g, float *vol, const float *projs, const double *eul, int n,
                     int iterations, double lambda, int update_nb, int first_index,
                     int startY, int endY, double *err, float *errvol)
{
    for (int it = 0; it < iterations; it++)
        if (osart_iteration_f32(g, vol, projs, eul, n, NULL, n, lambda, update_nb, first_index, startY, endY,
                                err + 5 * it, errvol))
            return -1;
    return 0;
}

/* OSSART over the available images `avail` (nav of them). */
int or_ossart_avail(const or_geom *g, float *vol, const float *projs, const double *eul, int n,
                    const int *avail, int nav, int iterations, double lambda, int update_nb, int first_index,
                    int startY, int endY, double *err)
{
    for (int it = 0; it < iterations; it++)
        if (osart_iteration_f32(g, vol, projs, eul, n, avail, nav, lambda, update_nb, first_index, startY, endY,
                                err + 5 * it, NULL))
            return -1;
    return 0;
}

int or_ossart_f64(const or_geom *g, double *vol, const double *projs, const double *eul, int n,
                  int iterations, double lambda, int update_nb, int first_index,
                  int startY, int endY, double *err)
{
    for (int it = 0; it < iterations; it++)
        if (osart_iteration_f64(g, vol, projs, eul, n, NULL, n, lambda, update_nb, first_index, startY, endY,
                                err + 5 * it))
            return -1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* WBP: Radermacher weighting mask, FFT filter                               */
/* ------------------------------------------------------------------------- */

/* FFTWeighting.createWeightingMask (FFTWeighting.java:31-65).  tilt/axis in degrees. */
void or_weighting_mask(int sx, int sy, const double *tilt_deg, const double *axis_deg, int n,
                       double diameter, float *mask)
{
    double cx = sx / 2.0, cy = sy / 2.0;
    double *sint = (double *)malloc(sizeof(double) * n);
    double *cosp = (double *)malloc(sizeof(double) * n);
    double *sinp = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) {
        sint[i] = sin(java8_to_radians(tilt_deg[i]));
        cosp[i] = cos(java8_to_radians(axis_deg[i]));
        sinp[i] = sin(java8_to_radians(axis_deg[i]));
    }
    for (int j = 0; j < sy; j++) {
        long jj = (long)j * sx;
        double cj = (j > cy) ? j - sy : j;
        for (int i = 0; i < sx; i++) {
            double ci = (i > cx) ? i - sx : i;
            double weight = 0;
            for (int w = 0; w < n; w++) {
                double zj = (ci * sint[w] * cosp[w] + cj * sint[w] * sinp[w]) / sx;
                double value = diameter * M_PI * zj;
                if (value == 0) value = 1;
                else value = sin(value) / value;
                weight += value;
            }
            if (weight < 0.005) weight = (double)0.005f;
            weight *= diameter;
            mask[jj + i] = (float)(1 / weight);
        }
    }
    free(sint); free(cosp); free(sinp);
}

/* In-place complex DFT along a strided line, double precision.  sign=-1 forward.
 * Radix-2 for powers of two, direct O(n^2) summation otherwise. */
static void dft_line(double *re, double *im, int n, int sign)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            for (int i = 0; i < n; i += len) {
                for (int k = 0; k < len / 2; k++) {
                    double ang = sign * 2.0 * M_PI * k / len;
                    double wr = cos(ang), wi = sin(ang);
                    double ur = re[i + k], ui = im[i + k];
                    double xr = re[i + k + len / 2], xi = im[i + k + len / 2];
                    double vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
                    re[i + k] = ur + vr; im[i + k] = ui + vi;
                    re[i + k + len / 2] = ur - vr; im[i + k + len / 2] = ui - vi;
                }
            }
        }
        return;
    }
    double *tr = (double *)malloc(sizeof(double) * n), *ti = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) {
        double sr = 0, si = 0;
        for (int t = 0; t < n; t++) {
            long m = ((long)k * t) % n;
            double ang = sign * 2.0 * M_PI * (double)m / n;
            sr += re[t] * cos(ang) - im[t] * sin(ang);
            si += re[t] * sin(ang) + im[t] * cos(ang);
        }
        tr[k] = sr; ti[k] = si;
    }
    memcpy(re, tr, sizeof(double) * n); memcpy(im, ti, sizeof(double) * n);
    free(tr); free(ti);
}

/* 2D complex DFT of a sy x sx row-major array, in place. */
void or_fft2(double *re, double *im, int sx, int sy, int sign)
{
    for (int r = 0; r < sy; r++) dft_line(re + (long)r * sx, im + (long)r * sx, sx, sign);
    double *cr = (double *)malloc(sizeof(double) * sy), *ci = (double *)malloc(sizeof(double) * sy);
    for (int c = 0; c < sx; c++) {
        for (int r = 0; r < sy; r++) { cr[r] = re[(long)r * sx + c]; ci[r] = im[(long)r * sx + c]; }
        dft_line(cr, ci, sy, sign);
        for (int r = 0; r < sy; r++) { re[(long)r * sx + c] = cr[r]; im[(long)r * sx + c] = ci[r]; }
    }
    free(cr); free(ci);
}

/* FFTWeighting.weighting (:87-112) in double: real(ifft2(mask * fft2(img))) with
 * ifft2 scaled by 1/(sx*sy).  img is overwritten (float in, float out). */
void or_weighting(float *img, const float *mask, int sx, int sy)
{
    long n = (long)sx * sy;
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)calloc(n, sizeof(double));
    for (long k = 0; k < n; k++) re[k] = img[k];
    or_fft2(re, im, sx, sy, -1);
    for (long k = 0; k < n; k++) { re[k] *= mask[k]; im[k] *= mask[k]; }
    or_fft2(re, im, sx, sy, +1);
    for (long k = 0; k < n; k++) img[k] = (float)(re[k] / (double)n);
    free(re); free(im);
}

void or_weighting_f64(double *img, const double *mask, int sx, int sy)
{
    long n = (long)sx * sy;
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)calloc(n, sizeof(double));
    for (long k = 0; k < n; k++) re[k] = img[k];
    or_fft2(re, im, sx, sy, -1);
    for (long k = 0; k < n; k++) { re[k] *= mask[k]; im[k] *= mask[k]; }
    or_fft2(re, im, sx, sy, +1);
    for (long k = 0; k < n; k++) img[k] = re[k] / (double)n;
    free(re); free(im);
}

/* TomoReconstruction2.WBP(ts, projector, params, startY, endY) (:499-538) with
 * update = 0: each projection is weighted (if mask != NULL) and backprojected
 * on its own.  vol must be zero on entry for a plain WBP. */
void or_wbp(const or_geom *g, float *vol, const float *projs, const double *eul, int n,
            const float *mask, int update_nb, int startY, int endY)
{
    const long npx = (long)g->pw * g->ph;
    float *buf = (float *)malloc(sizeof(float) * npx * (update_nb > 0 ? update_nb : 1));
    int *idxs = (int *)malloc(sizeof(int) * (update_nb > 0 ? update_nb : 1));
    int index = 0;
    for (int i = 0; i < n; i++) {
        memcpy(buf + (long)index * npx, projs + (size_t)i * npx, sizeof(float) * npx);
        idxs[index] = i;
        index++;
        if (index >= update_nb || i == n - 1) {
            for (int p = 0; p < index; p++) {
                if (mask) or_weighting(buf + (long)p * npx, mask, g->pw, g->ph);
                or_backproject_one(g, vol, buf + (long)p * npx, eul + (size_t)idxs[p] * 9, startY, endY);
            }
            index = 0;
        }
    }
    free(buf); free(idxs);
}

void or_wbp_f64(const or_geom *g, double *vol, const double *projs, const double *eul, int n,
                const double *mask, int startY, int endY)
{
    const long npx = (long)g->pw * g->ph;
    double *buf = (double *)malloc(sizeof(double) * npx);
    for (int i = 0; i < n; i++) {
        memcpy(buf, projs + (size_t)i * npx, sizeof(double) * npx);
        if (mask) or_weighting_f64(buf, mask, g->pw, g->ph);
        or_backproject_one_f64(g, vol, buf, eul + (size_t)i * 9, startY, endY);
    }
    free(buf);
}

/* ------------------------------------------------------------------------- */
/* Synthetic data: ellipsoid phantom, analytic FP64 projections, noise       */
/* ------------------------------------------------------------------------- */

static inline uint64_t splitmix64_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t v) { return (double)(v >> 11) * 0x1.0p-53; }

/* Ellipsoid record: 16 doubles = density, cx, cy, cz (volume-centred voxel
 * units), ax, ay, az (semi-axes), R[9] (rotation, body axes as rows). */
void or_ellipsoids(uint64_t seed, int k, int W, int H, int D, double *out)
{
    uint64_t s = seed;
    double L = W < H ? W : H;
    if (D < L) L = D;
    for (int e = 0; e < k; e++) {
        double *o = out + 16 * e;
        o[0] = 0.1 + 0.9 * u01(splitmix64_next(&s));
        double px, py, pz;
        do {
            px = 2 * u01(splitmix64_next(&s)) - 1;
            py = 2 * u01(splitmix64_next(&s)) - 1;
            pz = 2 * u01(splitmix64_next(&s)) - 1;
        } while (px * px + py * py + pz * pz > 1.0);
        o[1] = px * 0.4 * W; o[2] = py * 0.4 * H; o[3] = pz * 0.4 * D;
        for (int a = 0; a < 3; a++) o[4 + a] = (0.03 + 0.22 * u01(splitmix64_next(&s))) * L;
        /* Shoemake uniform quaternion */
        double u1 = u01(splitmix64_next(&s)), u2 = u01(splitmix64_next(&s)), u3 = u01(splitmix64_next(&s));
        double a = sqrt(1 - u1), b = sqrt(u1);
        double qw = a * sin(2 * M_PI * u2), qx = a * cos(2 * M_PI * u2);
        double qy = b * sin(2 * M_PI * u3), qz = b * cos(2 * M_PI * u3);
        double *R = o + 7;
        R[0] = 1 - 2 * (qy * qy + qz * qz); R[1] = 2 * (qx * qy - qz * qw); R[2] = 2 * (qx * qz + qy * qw);
        R[3] = 2 * (qx * qy + qz * qw); R[4] = 1 - 2 * (qx * qx + qz * qz); R[5] = 2 * (qy * qz - qx * qw);
        R[6] = 2 * (qx * qz - qy * qw); R[7] = 2 * (qy * qz + qx * qw); R[8] = 1 - 2 * (qx * qx + qy * qy);
    }
}

/* Line integral of the phantom along p + t*d (d unit), FP64. */
static double phantom_line(const double *ell, int k, const double *p, const double *d)
{
    double sum = 0;
    for (int e = 0; e < k; e++) {
        const double *o = ell + 16 * e;
        const double *R = o + 7;
        double q[3] = {p[0] - o[1], p[1] - o[2], p[2] - o[3]};
        double A = 0, B = 0, C = 0;
        for (int a = 0; a < 3; a++) {
            double qa = R[3 * a] * q[0] + R[3 * a + 1] * q[1] + R[3 * a + 2] * q[2];
            double da = R[3 * a] * d[0] + R[3 * a + 1] * d[1] + R[3 * a + 2] * d[2];
            double inv = 1.0 / (o[4 + a] * o[4 + a]);
            A += da * da * inv;
            B += 2 * da * qa * inv;
            C += qa * qa * inv;
        }
        C -= 1;
        double disc = B * B - 4 * A * C;
        if (disc > 0) sum += o[0] * sqrt(disc) / A;
    }
    return sum;
}

/* Analytic projections in TomoJ's geometry: detector pixel (i,j) of image p is
 * the line r = u*(i-pc0) + v*(j-pc1) + t*w, u,v,w the rows of E_p, in
 * volume-centred voxel units (scale 1).  out: n x ph x pw doubles. */
void or_phantom_project(const double *ell, int k, const double *eul, int n, int pw, int ph,
                        double pc0, double pc1, double *out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int pj = 0; pj < n * ph; pj++) {
        int p = pj / ph, j = pj % ph;
        const double *E = eul + 9 * p;
        for (int i = 0; i < pw; i++) {
            double a = i - pc0, b = j - pc1;
            double pt[3] = {E[0] * a + E[3] * b, E[1] * a + E[4] * b, E[2] * a + E[5] * b};
            double d[3] = {E[6], E[7], E[8]};
            out[((long)p * ph + j) * pw + i] = phantom_line(ell, k, pt, d);
        }
    }
}

/* Counter-based Gaussian noise: value n uses mix64(seed ^ 2n) and
 * mix64(seed ^ (2n+1)) through Box-Muller. */
void or_gaussian_noise(uint64_t seed, long count, double *out)
{
#pragma omp parallel for
    for (long n = 0; n < count; n++) {
        double u1 = ((double)((mix64(seed ^ (uint64_t)(2 * n)) >> 11) + 1)) * 0x1.0p-53;
        double u2 = u01(mix64(seed ^ (uint64_t)(2 * n + 1)));
        out[n] = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
}

/* Voxelised phantom (centre sampling), [z][y][x]. */
void or_phantom_volume(const double *ell, int k, int W, int H, int D, double vcx, double vcy,
                       double vcz, float *vol)
{
#pragma omp parallel for
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                double p[3] = {x - vcx, y - vcy, z - vcz}, v = 0;
                for (int e = 0; e < k; e++) {
                    const double *o = ell + 16 * e, *R = o + 7;
                    double q[3] = {p[0] - o[1], p[1] - o[2], p[2] - o[3]}, r = 0;
                    for (int a = 0; a < 3; a++) {
                        double qa = R[3 * a] * q[0] + R[3 * a + 1] * q[1] + R[3 * a + 2] * q[2];
                        r += qa * qa / (o[4 + a] * o[4 + a]);
                    }
                    if (r <= 1.0) v += o[0];
                }
                vol[((long)z * H + y) * W + x] = (float)v;
            }
}

int or_num_procs(void) { return omp_get_num_procs(); }

/* ---------------------------------------------------------------------------
 * FISTA extrapolation, FistaProjector3D.endOfIteration
 * (tomography/projectors/FistaProjector3D.java:43-59): t <- (1 + sqrt(1 + 4t^2))/2,
 * factor = (t_prev - 1)/t, and per voxel val = factor * (x - prev) with the
 * difference of the two floats taken in float (getPixel returns float), then
 * x += (float)val unless val is NaN (TomoReconstruction2.addToPixel :365-370).
 * *t carries the state (0 before the first iteration, FistaProjector3D.java:8).
 * ------------------------------------------------------------------------- */
void or_fista_step(float *vol, const float *prev, long n, double *t)
{
    const double t_1 = *t;
    const double tn = (1 + sqrt(1 + 4 * t_1 * t_1)) / 2;
    const double factor = (t_1 - 1) / tn;
    *t = tn;
    for (long i = 0; i < n; ++i) {
        const float d = vol[i] - prev[i];
        const double val = factor * (double)d;
        if (val != val) continue;
        vol[i] = vol[i] + (float)val;
    }
}

/* ---------------------------------------------------------------------------
 * Projection supply with the identity alignment: TiltSeries.getPixels(i)
 * (TiltSeries.java:678-687) -> AffineAlignment.applyTransformOnImage
 * (AffineAlignment.java:136-156) normalises the image by the statistics of a
 * tilt-dependent ROI (TiltSeries.getImageStatistics :809-835): a rectangle of
 * roisx = max(round(W cos(tilt)), minSize) columns centred by integer
 * division, all rows.  Statistics as ImageJ 1.53a FloatStatistics (third
 * party, restated): double sums over the ROI row by row, mean = sum/n,
 * stdDev = sqrt(((n*sum2 - sum*sum)/n)/(n-1)) (0 when not positive).
 * ------------------------------------------------------------------------- */
void or_roi_rect(int width, double tilt_deg, int min_size, int *x0, int *x1)
{
    const double rad = tilt_deg / 180.0 * M_PI;
    long r = (long)floor(width * cos(rad) + 0.5); /* Math.round */
    int roisx = (int)(r > min_size ? r : min_size);
    const double cx = (width - 1.0) / 2;
    const double p0 = (double)((width - roisx) / 2) - cx; /* int division, then - cx */
    *x0 = (int)(p0 + cx);
    *x1 = (int)((p0 + roisx) + cx);
}

void or_roi_stats(const float *img, int width, int height, int x0, int x1, double *mean, double *sd)
{
    double n = 0, sum = 0, sum2 = 0;
    if (x0 < 0) x0 = 0;
    if (x1 > width) x1 = width;
    for (int y = 0; y < height; ++y)
        for (int x = x0; x < x1; ++x) {
            const double v = img[(long)y * width + x];
            sum += v;
            sum2 += v * v;
            n += 1;
        }
    *mean = n > 0 ? sum / n : 0.0;
    double s = 0.0;
    if (n > 0) {
        s = (n * sum2 - sum * sum) / n;
        s = (s > 0.0) ? sqrt(s / (n - 1.0)) : 0.0;
    }
    *sd = s;
}

void or_normalize(float *img, long npx, double mean, double sd)
{
    for (long i = 0; i < npx; ++i) img[i] = (float)((img[i] - mean) / sd);
}

/* ---------------------------------------------------------------------------
 * Aligned projection supply: AffineAlignment.applyTransformOnImage for a
 * non-identity transform (tomography/align/AffineAlignment.java:157-254) and
 * the ROI of TiltSeries.getImageStatistics (tomography/TiltSeries.java:809-835).
 * Matrices are 6 doubles in java.awt.geom.AffineTransform.getMatrix order:
 * m00, m10, m01, m11, m02, m12.
 * ------------------------------------------------------------------------- */

/* java.awt.geom.AffineTransform.createInverse: the inverse per transform state
 * (identity / translation, scale (+translation), pure shear (+translation),
 * general).  Returns 0 for a non-invertible transform. */
int or_affine_inverse(const double *m, double *r)
{
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5];
    int tr = (e != 0.0 || f != 0.0);
    if (c == 0.0 && b == 0.0) {
        if (a == 1.0 && d == 1.0) {
            r[0] = 1.0; r[1] = 0.0; r[2] = 0.0; r[3] = 1.0;
            r[4] = tr ? -e : 0.0;
            r[5] = tr ? -f : 0.0;
            return 1;
        }
        if (a == 0.0 || d == 0.0) return 0;
        r[0] = 1.0 / a; r[1] = 0.0; r[2] = 0.0; r[3] = 1.0 / d;
        r[4] = tr ? -e / a : 0.0;
        r[5] = tr ? -f / d : 0.0;
        return 1;
    }
    if (a == 0.0 && d == 0.0) {
        if (c == 0.0 || b == 0.0) return 0;
        r[0] = 0.0; r[1] = 1.0 / c; r[2] = 1.0 / b; r[3] = 0.0;
        r[4] = tr ? -f / b : 0.0;
        r[5] = tr ? -e / c : 0.0;
        return 1;
    }
    {
        double det = a * d - c * b;
        if (fabs(det) <= 4.9406564584124654e-324) return 0;
        r[0] = d / det; r[1] = -b / det; r[2] = -c / det; r[3] = a / det;
        r[4] = tr ? (c * f - d * e) / det : 0.0;
        r[5] = tr ? (b * e - a * f) / det : 0.0;
    }
    return 1;
}

/* One aligned image (AffineAlignment.java:157-254): inv is the inverse
 * transform, centres w/2, h/2; fill 0 none, 1 average, 2 NaN. */
void or_affine_resample(const float *src, int w, int h, const double *inv, int normalize, int fill, double mean,
                        double sd, float *dst)
{
    double cx = w / 2.0, cy = h / 2.0;
    for (int j = 0; j < h; ++j) {
        double yc = j - cy;
        for (int i = 0; i < w; ++i) {
            double xc = i - cx;
            double xx = (inv[0] * xc + inv[2] * yc + inv[4]) + cx;
            double yy = (inv[1] * xc + inv[3] * yc + inv[5]) + cy;
            int ix0 = (int)xx, iy0 = (int)yy;
            double dx0 = xx - ix0, dy0 = yy - iy0;
            float pix;
            if (ix0 >= 0 && ix0 < w && iy0 >= 0 && iy0 < h) {
                if (ix0 == w - 1 || iy0 == h - 1) {
                    pix = src[ix0 + (long)iy0 * w];
                } else {
                    double fac4 = dx0 * dy0;
                    long pos = ix0 + (long)iy0 * w;
                    double v1 = src[pos], v3 = src[pos + w], v2 = src[pos + 1], v4 = src[pos + 1 + w];
                    pix = (float)(v1 * (1 - dx0 - dy0 + fac4) + v2 * (dx0 - fac4) + v3 * (dy0 - fac4) + v4 * fac4);
                }
                if (normalize) pix = (float)((pix - mean) / sd);
            } else if (fill == 1) {
                pix = normalize ? 0.0f : (float)mean;
            } else if (fill == 2) {
                pix = NAN;
            } else {
                pix = 0.0f;
            }
            dst[(long)j * w + i] = pix;
        }
    }
}

/* The ROI polygon of getImageStatistics mapped by T, for a transform without
 * rotation or shear: the rectangle [x0, x1) x [y0, y1) of its int corners. */
void or_roi_rect_affine(int width, int height, double tilt_deg, int min_size, const double *m, int *r4)
{
    int x0, x1;
    double rad = tilt_deg / 180.0 * M_PI;
    long r = (long)floor(width * cos(rad) + 0.5);
    int roisx = (int)(r > min_size ? r : min_size);
    double cx = (width - 1.0) / 2, cy = (height - 1.0) / 2;
    double pts[8];
    int xs[4], ys[4];
    pts[0] = (double)((width - roisx) / 2) - cx;
    pts[1] = -cy;
    pts[2] = pts[0] + roisx;
    pts[3] = pts[1];
    pts[4] = pts[2];
    pts[5] = height - cy;
    pts[6] = pts[0];
    pts[7] = pts[5];
    for (int k = 0; k < 4; ++k) {
        xs[k] = (int)((m[0] * pts[2 * k] + m[2] * pts[2 * k + 1] + m[4]) + cx);
        ys[k] = (int)((m[1] * pts[2 * k] + m[3] * pts[2 * k + 1] + m[5]) + cy);
    }
    x0 = xs[0]; x1 = xs[0];
    r4[2] = ys[0]; r4[3] = ys[0];
    for (int k = 1; k < 4; ++k) {
        if (xs[k] < x0) x0 = xs[k];
        if (xs[k] > x1) x1 = xs[k];
        if (ys[k] < r4[2]) r4[2] = ys[k];
        if (ys[k] > r4[3]) r4[3] = ys[k];
    }
    r4[0] = x0;
    r4[1] = x1;
}

/* Mean and sample standard deviation over the rectangle [x0, x1) x [y0, y1)
 * (clipped), as or_roi_stats. */
void or_rect_stats(const float *img, int width, int height, const int *r4, double *mean, double *sd)
{
    double n = 0, sum = 0, sum2 = 0, s = 0.0;
    int x0 = r4[0] < 0 ? 0 : r4[0], x1 = r4[1] > width ? width : r4[1];
    int y0 = r4[2] < 0 ? 0 : r4[2], y1 = r4[3] > height ? height : r4[3];
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            double v = img[(long)y * width + x];
            sum += v;
            sum2 += v * v;
            n += 1;
        }
    *mean = n > 0 ? sum / n : 0.0;
    if (n > 0) {
        s = (n * sum2 - sum * sum) / n;
        s = (s > 0.0) ? sqrt(s / (n - 1.0)) : 0.0;
    }
    *sd = s;
}

/* ---------------------------------------------------------------------------
 * Total-variation denoising, TotalVariationDenoising_3D.denoise3D
 * (src/Filters/src/main/java/fr/curie/filters/TotalVariationDenoising_3D.java:84-202):
 * the end-of-iteration step of TVMProjector3D (src/TomoJ_AdvancedReconstruction/
 * src/main/java/fr/curie/tomoj/tomography/projectors/TVMProjector3D.java:135-157).
 * Volume f in ImageStack layout [z][y][x], replaced by the result.  u starts as
 * a copy of f and p (three floats per voxel, interleaved x, y, z as pstack
 * holds them) at zero; each step runs updateP over every voxel, then updateU.
 * Float arithmetic in Java's order, nothing fused:
 *   updateP (:115-156): forward differences duy, dux, duz, 0 on the last row,
 *     column or slice; dd1 = 1 + dt/theta/g * |(float)sqrt(duy^2 + dux^2 + duz^2)|;
 *     p_c = (p_c - dt/theta * du_c) / dd1.
 *   updateU (:158-202): div = p_x(x) at x = 0, -p_x(x) at x = W-1 (the voxel's
 *     own component, not its left neighbour's), p_x(x) - p_x(x-1) between;
 *     then the same for y and z added on; u = f - theta * div.
 * ------------------------------------------------------------------------- */
void or_tv_denoise3d(float *f, int W, int H, int D, float theta, float g, float dt, int iterations)
{
    const long plane = (long)W * H, n = plane * D;
    float *u = (float *)malloc(sizeof(float) * n);
    float *p = (float *)calloc((size_t)n * 3, sizeof(float));
    if (!u || !p) {
        free(u);
        free(p);
        return;
    }
    memcpy(u, f, sizeof(float) * n);
    for (int it = 0; it < iterations; ++it) {
        /* updateP: slice z reads slice min(z+1, D-1) for the z difference */
        for (int z = 0; z < D; ++z) {
            const float *us = u + z * plane;
            const float *u1 = u + (z + 1 < D ? z + 1 : D - 1) * plane;
            float *ps = p + 3 * z * plane;
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    const long i = (long)y * W + x;
                    const float duy = (y < H - 1) ? us[i + W] - us[i] : 0.0f;
                    const float dux = (x < W - 1) ? us[i + 1] - us[i] : 0.0f;
                    const float duz = (z < D - 1) ? u1[i] - us[i] : 0.0f;
                    const float dd1 = 1 + dt / theta / g * fabsf((float)sqrt((double)(duy * duy + dux * dux + duz * duz)));
                    ps[3 * i] = (ps[3 * i] - dt / theta * dux) / dd1;
                    ps[3 * i + 1] = (ps[3 * i + 1] - dt / theta * duy) / dd1;
                    ps[3 * i + 2] = (ps[3 * i + 2] - dt / theta * duz) / dd1;
                }
        }
        /* updateU: slice z reads p of slice max(z-1, 0) for the z divergence */
        for (int z = 0; z < D; ++z) {
            const float *ps = p + 3 * z * plane;
            const float *p1 = p + 3 * (z > 0 ? z - 1 : 0) * plane;
            const float *fs = f + z * plane;
            float *us = u + z * plane;
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    const long i = (long)y * W + x;
                    float div;
                    if (x == 0) div = ps[3 * i];
                    else if (x == W - 1) div = -ps[3 * i];
                    else div = ps[3 * i] - ps[3 * (i - 1)];
                    if (y == 0) div += ps[3 * i + 1];
                    else if (y == H - 1) div += -ps[3 * i + 1];
                    else div += ps[3 * i + 1] - ps[3 * (i - W) + 1];
                    if (z == 0) div += ps[3 * i + 2];
                    else if (z == D - 1) div += -ps[3 * i + 2];
                    else div += ps[3 * i + 2] - p1[3 * i + 2];
                    us[i] = fs[i] - theta * div;
                }
        }
    }
    memcpy(f, u, sizeof(float) * n);
    free(u);
    free(p);
}

/* The 2-D form, TotalVariationDenoising_3D.denoise2D (:206-260), on one W x H
 * image in place: forward differences of the whole image first (du holds the
 * y difference in its first half and the x difference in its second, and keeps
 * the previous step's value where no difference is formed), then p, then the
 * divergence, then u.  A volume whose slices are all equal has duz = 0 and
 * p_z = 0 throughout, so denoise3D reduces to this on each slice. */
void or_tv_denoise2d(float *img, int w, int h, float theta, float g, float dt, int iterations)
{
    const long n = (long)w * h;
    float *u = (float *)malloc(sizeof(float) * n);
    float *p = (float *)calloc((size_t)n * 2, sizeof(float));
    float *du = (float *)calloc((size_t)n * 2, sizeof(float));
    float *divp = (float *)calloc((size_t)n, sizeof(float));
    if (!u || !p || !du || !divp) {
        free(u);
        free(p);
        free(du);
        free(divp);
        return;
    }
    memcpy(u, img, sizeof(float) * n);
    for (int it = 0; it < iterations; ++it) {
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < w; ++i) {
                if (i < w - 1) du[i + (long)w * (j + h)] = u[i + 1 + (long)w * j] - u[i + (long)w * j];
                if (j < h - 1) du[i + (long)w * j] = u[i + (long)w * (j + 1)] - u[i + (long)w * j];
            }
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < w; ++i) {
                const float du1 = du[i + (long)w * j], du2 = du[i + (long)w * (j + h)];
                const float dd1 = 1 + dt / theta / g * fabsf((float)sqrt((double)(du1 * du1 + du2 * du2)));
                p[i + (long)w * j] = (p[i + (long)w * j] - dt / theta * du[i + (long)w * j]) / dd1;
                p[i + (long)w * (j + h)] = (p[i + (long)w * (j + h)] - dt / theta * du[i + (long)w * (j + h)]) / dd1;
            }
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < w; ++i) {
                if (i == 0) divp[(long)w * j] = p[(long)w * (j + h)];
                else if (i == w - 1) divp[w - 1 + (long)w * j] = -p[w - 1 + (long)w * (j + h)];
                else divp[i + (long)w * j] = p[i + (long)w * (j + h)] - p[i - 1 + (long)w * (j + h)];
                if (j == 0) divp[i] += p[i];
                else if (j == h - 1) divp[i + (long)w * (h - 1)] += -p[i + (long)w * (h - 1)];
                else divp[i + (long)w * j] += p[i + (long)w * j] - p[i + (long)w * (j - 1)];
            }
        for (long k = 0; k < n; ++k) u[k] = img[k] - theta * divp[k];
    }
    memcpy(img, u, sizeof(float) * n);
    free(u);
    free(p);
    free(du);
    free(divp);
}

/* ---------------------------------------------------------------------------
 * Bayesian regularisation, BayesianVoxelProjector3D (src/TomoJ_AdvancedReconstruction/
 * src/main/java/fr/curie/tomoj/tomography/projectors/BayesianVoxelProjector3D.java).
 * Volumes in ImageStack layout [z][y][x].
 *
 * or_bayes_convolve: startOfIteration (:134-137) -> convolve(rec, kernel,
 * distanceVolume, -lambda) (:175-198).  For each voxel the 27 values of
 * TomoReconstruction2.getCubicNeightborhoodWithClamp(x, y, z, 1) (TomoReconstruction2.java:317-339;
 * despite its name it reads getPixel(x+i, y+j, z+k), 0 outside the volume, :241-246),
 * i over x outermost and k over z innermost, each times its kernel entry (:31-33,
 * the separable [1,-2,1] stencil) in float, summed in double in that order;
 * dist = (float)(total * factor) (putPixel :351-353).
 * ------------------------------------------------------------------------- */
static const float or_bayes_kernel[27] = {1, -2, 1, -2, 4, -2, 1, -2, 1,
                                          -2, 4, -2, 4, -8, 4, -2, 4, -2,
                                          1, -2, 1, -2, 4, -2, 1, -2, 1};

static float or_vox(const float *v, int W, int H, int D, int x, int y, int z)
{
    if (x < 0 || x >= W || y < 0 || y >= H || z < 0 || z >= D) return 0.0f;
    return v[((long)z * H + y) * W + x];
}

void or_bayes_convolve(const float *vol, float *dist, int W, int H, int D, double factor)
{
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                double total = 0;
                int idx = 0;
                for (int i = -1; i <= 1; ++i)
                    for (int j = -1; j <= 1; ++j)
                        for (int k = -1; k <= 1; ++k) {
                            const float prod = or_vox(vol, W, H, D, x + i, y + j, z + k) * or_bayes_kernel[idx++];
                            total += prod;
                        }
                dist[((long)z * H + y) * W + x] = (float)(total * factor);
            }
}

/* or_bayes_end: endOfIteration (:121-132): mult(distanceVolume, alpha) (:159-173,
 * ImageJ FloatProcessor.multiply: the factor rounded to float, then a float
 * product), rec.operation(distanceVolume, ADD) (TomoReconstruction2.java:386-403,
 * FloatBlitter ADD: a float sum), then with positivity
 * rec.applyPositivityConstraint() (:405-412, FloatBlitter MAX against 0: a value
 * is replaced only where 0 > value, so NaN and -0 stay).  ImageJ 1.53a's
 * FloatProcessor / FloatBlitter are third party, restated from their published
 * source. */
void or_bayes_end(float *vol, float *dist, long n, double alpha, int positivity)
{
    const float c = (float)alpha;
    for (long i = 0; i < n; ++i) {
        dist[i] = dist[i] * c;
        float v = dist[i] + vol[i];
        if (positivity && 0.0f > v) v = 0.0f;
        vol[i] = v;
    }
}

/* ---------------------------------------------------------------------------
 * Hysteresis threshold of the projection supply, TiltSeries.applyThreshold
 * (src/TomoJ_Base/.../tomography/TiltSeries.java:720-806), applied by getPixels
 * (:678-686) after the alignment.  One image in place.
 *
 * Statistics as ImageJ 1.53a's FloatStatistics (third party, restated from its
 * published source): values v with -FLT_MAX <= v <= FLT_MAX (NaN and +-Inf
 * left out), histMin/histMax = their min/max, 256 bins, bin (int)(scale *
 * (v - histMin)) with scale = 256 / (histMax - histMin) (Java's (int) of NaN is
 * 0: a constant image fills bin 0), clamped to 255; mean = sequential double
 * sum / count.  Levels: th = (last bin i whose running count is still below
 * (int)(W*H*fraction)) / 256 * (histMax - histMin) + histMin, for the
 * threshold and the hysteresis fraction.  Pixels below th1 take the fill value
 * (NaN, 0, or (float) mean); pixels in [th1, th2) are queued, and a queued
 * pixel takes the fill value once a neighbour holds it (NaN in NaN mode; the
 * neighbours are index-1 and index+1, wrapping across rows as TomoJ's test
 * does, and index-W, index+W), sweeping the queue backwards until a sweep
 * changes nothing.
 * ------------------------------------------------------------------------- */
static int java_d2i(double x)
{
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return -2147483647 - 1;
    return (int)x;
}

void or_threshold(float *res, int width, int height, double threshold, double hysteresis, int fill)
{
    const long len = (long)width * height;
    double mn = DBL_MAX, mx = -DBL_MAX, sum = 0;
    long cnt = 0;
    int hist[256];
    memset(hist, 0, sizeof hist);
    for (long i = 0; i < len; i++) {
        const double v = res[i];
        if (v >= -FLT_MAX && v <= FLT_MAX) {
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
    }
    const double scale = 256 / (mx - mn);
    for (long i = 0; i < len; i++) {
        const double v = res[i];
        if (v >= -FLT_MAX && v <= FLT_MAX && v >= mn && v <= mx) {
            cnt++;
            sum += v;
            int index = java_d2i(scale * (v - mn));
            if (index >= 256) index = 255;
            hist[index]++;
        }
    }
    const double mean = sum / cnt;
    const int pixnb1 = java_d2i((double)width * height * threshold);
    const int pixnb2 = java_d2i((double)width * height * hysteresis);
    double th1 = 0, th2 = 0;
    long run = 0;
    for (int i = 0; i < 256; i++) {
        run += hist[i];
        if (run < pixnb1) th1 = i;
        if (run < pixnb2) th2 = i;
    }
    th1 /= 256;
    th1 = th1 * (mx - mn) + mn;
    th2 /= 256;
    th2 = th2 * (mx - mn) + mn;
    float value = NAN;
    if (fill == 1) value = (float)mean;
    if (fill == 0) value = 0;
    const int nan_mode = (value != value);
    long *q = (long *)malloc(sizeof(long) * (len > 0 ? len : 1));
    long nq = 0, count = 0;
    for (long i = 0; i < len; i++) {
        if (res[i] < th1) res[i] = value;
        else if (res[i] < th2) {
            count++;
            q[nq++] = i;
        }
    }
#define OR_MATCH(j) (nan_mode ? (res[j] != res[j]) : (res[j] == value))
    while (count > 0) {
        count = 0;
        for (long i = nq - 1; i >= 0; i--) {
            const long ind = q[i];
            int hit = 0;
            if (ind - 1 >= 0 && OR_MATCH(ind - 1)) hit = 1;
            else if (ind + 1 < len && OR_MATCH(ind + 1)) hit = 1;
            else if (ind - width >= 0 && OR_MATCH(ind - width)) hit = 1;
            else if (ind + width < len && OR_MATCH(ind + width)) hit = 1;
            if (hit) {
                count++;
                memmove(q + i, q + i + 1, sizeof(long) * (nq - i - 1));
                nq--;
                res[ind] = value;
            }
        }
    }
#undef OR_MATCH
    free(q);
}

/* ---------------------------------------------------------------------------
 * Statistics of TiltSeries.getImageStatistics' polygon ROI under a rotating or
 * shearing alignment (TiltSeries.java:809-835): the four int corners
 * (xs[k], ys[k]) of the mapped rectangle define an ImageJ PolygonRoi, and
 * FloatStatistics sums the pixels of its mask.  Mask restated from ImageJ's
 * documented rule (third party, not in the tree): a pixel belongs to the
 * polygon when its centre (x + 0.5, y + 0.5) lies inside by the even-odd
 * crossing test; the test's half-open convention decides centres exactly on
 * an edge.  Pixels outside the image are left out.  Sums row by row over the
 * bounding box, mean and sample deviation as or_rect_stats.
 * ------------------------------------------------------------------------- */
int or_poly_inside(const int *xs, const int *ys, double px, double py)
{
    int in = 0;
    for (int i = 0, j = 3; i < 4; j = i++) {
        const double xi = xs[i], yi = ys[i], xj = xs[j], yj = ys[j];
        if ((yi > py) != (yj > py) && px < (xj - xi) * (py - yi) / (yj - yi) + xi) in = !in;
    }
    return in;
}

void or_poly_stats(const float *img, int width, int height, const int *poly, double *mean, double *sd)
{
    const int *xs = poly, *ys = poly + 4;
    int x0 = xs[0], x1 = xs[0], y0 = ys[0], y1 = ys[0];
    for (int k = 1; k < 4; k++) {
        if (xs[k] < x0) x0 = xs[k];
        if (xs[k] > x1) x1 = xs[k];
        if (ys[k] < y0) y0 = ys[k];
        if (ys[k] > y1) y1 = ys[k];
    }
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > width) x1 = width;
    if (y1 > height) y1 = height;
    double n = 0, sum = 0, sum2 = 0, s = 0.0;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x)
            if (or_poly_inside(xs, ys, x + 0.5, y + 0.5)) {
                const double v = img[(long)y * width + x];
                sum += v;
                sum2 += v * v;
                n += 1;
            }
    *mean = n > 0 ? sum / n : 0.0;
    if (n > 0) {
        s = (n * sum2 - sum * sum) / n;
        s = (s > 0.0) ? sqrt(s / (n - 1.0)) : 0.0;
    }
    *sd = s;
}

/* The ROI polygon's int corners (xs[4], ys[4]) for an image of width x height
 * at tilt `tilt_deg` under the forward transform m (m00, m10, m01, m11, m02, m12),
 * TiltSeries.getImageStatistics (:809-828). */
void or_roi_poly_affine(int width, int height, double tilt_deg, int min_size, const double *m, int *poly)
{
    double rad = tilt_deg / 180.0 * M_PI;
    long r = (long)floor(width * cos(rad) + 0.5);
    int roisx = (int)(r > min_size ? r : min_size);
    double cx = (width - 1.0) / 2, cy = (height - 1.0) / 2;
    double pts[8];
    pts[0] = (double)((width - roisx) / 2) - cx;
    pts[1] = -cy;
    pts[2] = pts[0] + roisx;
    pts[3] = pts[1];
    pts[4] = pts[2];
    pts[5] = height - cy;
    pts[6] = pts[0];
    pts[7] = pts[5];
    for (int k = 0; k < 4; ++k) {
        poly[k] = (int)((m[0] * pts[2 * k] + m[2] * pts[2 * k + 1] + m[4]) + cx);
        poly[4 + k] = (int)((m[1] * pts[2 * k] + m[3] * pts[2 * k + 1] + m[5]) + cy);
    }
}

/* ---------------------------------------------------------------------------
 * BgART projector, BgARTVoxelProjector3D (src/TomoJ_AdvancedReconstruction/src/
 * main/java/fr/curie/tomoj/tomography/projectors/BgARTVoxelProjector3D.java).
 * Volumes in ImageStack layout [z][y][x].
 * ------------------------------------------------------------------------- */

/* Java's Float.compare order (Arrays.sort(float[])): -0.0 before 0.0, NaN last. */
static int java_float_key(float v)
{
    int32_t b;
    if (v != v) return 0x7fc00000;
    memcpy(&b, &v, 4);
    return b >= 0 ? b : (int32_t)(b ^ 0x7fffffff);
}

/* medianFilter3Dcross (:123-151): each voxel takes the median (sorted[3]) of
 * itself, the voxels of the slices before and after (0 beyond the stack) and
 * the pixels index -+1 and -+W of its slice image (index +-1 wraps across rows;
 * the tests are idx+1 < W*H, idx-1 > 0, idx-W > 0, idx+W < W*H, 0 otherwise),
 * all read from a copy of the volume. */
void or_median_cross(float *vol, int W, int H, int D)
{
    const long plane = (long)W * H, n = plane * D;
    float *tmp = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
    memcpy(tmp, vol, sizeof(float) * n);
    for (int z = 0; z < D; z++) {
        const float *pixs = tmp + z * plane;
        const float *pmin = z > 0 ? tmp + (z - 1) * plane : NULL;
        const float *pmax = z < D - 1 ? tmp + (z + 1) * plane : NULL;
        for (long i = 0; i < plane; i++) {
            float v[7];
            v[0] = pmin ? pmin[i] : 0.0f;
            v[1] = pmax ? pmax[i] : 0.0f;
            v[2] = pixs[i];
            v[3] = (i + 1 < plane) ? pixs[i + 1] : 0.0f;
            v[4] = (i - 1 > 0) ? pixs[i - 1] : 0.0f;
            v[5] = (i - W > 0) ? pixs[i - W] : 0.0f;
            v[6] = (i + W < plane) ? pixs[i + W] : 0.0f;
            for (int a = 1; a < 7; a++) {   /* insertion sort by Float.compare */
                float x = v[a];
                int b = a - 1;
                while (b >= 0 && java_float_key(v[b]) > java_float_key(x)) {
                    v[b + 1] = v[b];
                    b--;
                }
                v[b + 1] = x;
            }
            vol[z * plane + i] = v[3];
        }
    }
    free(tmp);
}

/* collapseFilter (:34-78): ImageJ StackStatistics of the volume (third party,
 * restated as FloatStatistics: values within +-FLT_MAX, min/max, 256 bins of
 * (int)(256/(max-min) * (v - min)) clamped to 255), the background value =
 * min + (first fullest bin) * binSize (or bg_value when it is not NaN), sigma
 * from the bins below it, and every voxel below mv + k sigma (dark
 * background; above mv - k sigma otherwise) set to the background.  Returns
 * the background value. */
double or_collapse(float *vol, long n, double k, int dark_bg, double bg_value)
{
    double mn = DBL_MAX, mx = -DBL_MAX;
    int hist[256];
    memset(hist, 0, sizeof hist);
    for (long i = 0; i < n; i++) {
        const double v = vol[i];
        if (v >= -FLT_MAX && v <= FLT_MAX) {
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
    }
    const double scale = 256 / (mx - mn);
    for (long i = 0; i < n; i++) {
        const double v = vol[i];
        if (v >= -FLT_MAX && v <= FLT_MAX && v >= mn && v <= mx) {
            int index = java_d2i(scale * (v - mn));
            if (index >= 256) index = 255;
            hist[index]++;
        }
    }
    const double bin = (mx - mn) / 256;
    double best = 0;
    int bi = 0;
    for (int i = 0; i < 256; i++)
        if (hist[i] > best) {
            best = hist[i];
            bi = i;
        }
    const double mv = (bg_value != bg_value) ? mn + bi * bin : bg_value;
    double sigma = 0;
    int count = 0;
    for (int i = 0; i < bi; i++) {
        sigma += (mn + i * bin - mv) * (mn + i * bin - mv) * hist[i];
        count += hist[i];
    }
    sigma = sqrt(sigma / count);
    const double th1 = mv - k * sigma, th2 = mv + k * sigma;
    for (long i = 0; i < n; i++) {
        const double v = vol[i];
        if (dark_bg && v < th2) vol[i] = (float)mv;
        if (!dark_bg && v > th1) vol[i] = (float)mv;
    }
    return mv;
}
