/*
 * tomoj_hip.h -- C ABI of the MI355X projector library (libtomoj_hip.so).
 *
 * The library stands where TomoJ's OpenCL projector stands today: behind the
 * Java `Projector` plugin interface
 *   src/TomoJ_Base/src/main/java/fr/curie/tomoj/tomography/projectors/Projector.java:14-159
 * and the GPU lifecycle that ReconstructionThreadManagerGPU drives on
 *   src/TomoJ_GPU/src/main/java/fr/curie/gpu/tomoj/tomography/projectors/ProjectorGPU.java.
 * Every entry point below names the reference method it replaces.
 *
 * Conventions
 *   - Plain C types only; host arrays are owned by the caller and copied during
 *     the call (never retained).  Device memory belongs to the tj_ctx.
 *   - Host volume layout is TomoJ's ImageStack layout: slice z, row y, x fastest
 *     (`float zyx[D][H][W]`).  Projections are `float [n][ph][pw]`.
 *   - Euler matrices are row-major 3x3 doubles, as returned by
 *     Alignment.getEulerMatrix (align/AffineAlignment.java:382-395).
 *   - Every call returns 0 on success or a negative TJ_E* code; the message of the
 *     last failure on the calling thread is returned by tj_last_error().
 *   - A tj_ctx is driven by one host thread at a time (as one Java thread drives
 *     one ProjectorGPU); tj_get_completion may be polled from another thread.
 *   - All calls are synchronous: when they return, the device work is done
 *     (TomoJ ends each step with clFinish).
 */
#ifndef TOMOJ_HIP_H
#define TOMOJ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TJ_OK 0
#define TJ_EINVAL -1     /* bad argument (sizes, indices, null pointers)          */
#define TJ_ENOMEM -2     /* device or host allocation failed / over HBM budget    */
#define TJ_EDEVICE -3    /* HIP runtime error                                     */
#define TJ_ESTATE -4     /* call not valid in the current state                  */
#define TJ_EUNSUP -5     /* geometry or option not supported                     */
#define TJ_ECANCELLED -6 /* tj_interrupt() was called during the loop             */
#define TJ_EIO -7        /* a file could not be read or written                   */

/* Storage type of the resident volume and projections.  The projector context
 * (tj_open) stores FP32 only.  The job engine (tj_batch_*) stores any of them:
 * FP32, BF16 and FP16 storage compute in FP32, FP64 storage computes in FP64
 * (the reference twin used to measure the others); residual norms are always
 * reduced in FP64. */
#define TJ_STORAGE_F32 0
#define TJ_STORAGE_BF16 1
#define TJ_STORAGE_F16 2
#define TJ_STORAGE_F64 3

typedef struct tj_ctx tj_ctx;

/* Projector state TomoJ keeps in Projector / VoxelProjector3D fields. */
typedef struct {
    int vol_w, vol_h, vol_d;  /* TomoReconstruction2 width, height, #slices             */
    int proj_w, proj_h;       /* TiltSeries width, height                                */
    double proj_center[4];    /* Projector.projectionCenters: [0..1] used by the forward
                                 projection (ts.getCenterX/Y), [2..3] by the
                                 backprojection (ts.getProjectionCenterX/Y)
                                 (Projector.java:52-58)                                  */
    double vol_center[3];     /* Projector.volumeCenters (Projector.java:60-65)         */
    double scale;             /* VoxelProjector3D.setScale (VoxelProjector3D.java:89-91) */
    int positivity;           /* setPositivityConstraint (VoxelProjector3D.java:102-104) */
    int long_object_comp;     /* setLongObjectCompensation (VoxelProjector3D.java:78-80) */
    int storage;              /* TJ_STORAGE_*                                            */
} tj_geometry;

/* Fills g with TomoJ's defaults for a vol_w x vol_h x vol_d reconstruction of a
 * proj_w x proj_h tilt series: centres (n-1)/2 (TiltSeries.java:161-170,
 * TomoReconstruction2.java:63-71), scale 1, no positivity, no long-object
 * compensation, FP32 storage. */
void tj_geometry_default(tj_geometry *g, int vol_w, int vol_h, int vol_d, int proj_w, int proj_h);

const char *tj_last_error(void);
const char *tj_version(void);

/* JSON description of HIP device `device` (name, arch, CUs, clock, LDS sizes,
 * HBM and L2 bytes) into buf.  No TomoJ counterpart (JOCLDeviceInfo plays this
 * role for OpenCL devices). */
int tj_device_info(int device, char *buf, int len);

/* ProjectorGPU.computeMaximumSizeForVolume (ProjectorGPU.java:162-166): the
 * bytes of volume a slab driver may place on `device`, min(0.4 x its global
 * memory, its largest allocation). */
int tj_max_volume_bytes(int device, size_t *bytes);

/* Copies `bytes` between two device buffers `repeats` times with a plain
 * streaming kernel and reports the average milliseconds per copy.  No TomoJ
 * counterpart: it is the known-size reference used to calibrate HBM byte
 * counters (FETCH_SIZE / WRITE_SIZE) and the achievable copy bandwidth. */
int tj_copy_probe(int device, size_t bytes, int repeats, double *ms);

/* Creates a projector context on HIP device `device` and allocates the resident
 * volume.  Replaces `new VoxelProjector3DGPU(...)` + ProjectorGPU.initForIterative
 * buffer allocation (ProjectorGPU.java:195-254).  The volume starts at zero. */
int tj_open(int device, const tj_geometry *g, tj_ctx **out);

/* Frees all device memory.  Replaces ProjectorGPU.releaseCL_Memory (ProjectorGPU.java:376-403). */
int tj_close(tj_ctx *ctx);

/* Makes n projections (already supplied, i.e. TiltSeries.getPixels(i) for each i,
 * TiltSeries.java:678-687) and their Euler matrices resident on the device.
 * `stack` is n x proj_h x proj_w, `euler` n x 9.  TomoJ re-uploads each image on
 * every addProjection (ProjectorGPU.java:411-438); here they are uploaded once. */
int tj_load_projections(tj_ctx *ctx, const float *stack, int n, const double *euler);

/* Replaces the resident Euler matrices only (alignment changed). */
int tj_set_eulers(tj_ctx *ctx, const double *euler, int n);

/* Copies volume rows [start_y,end_y) of the host volume `zyx` (full D x H x W
 * array) to the device.  Replaces the slab upload in ProjectorGPU.initForIterative
 * (ProjectorGPU.java:195-254) / copyInGPUBuffer (:317-330). */
int tj_set_volume(tj_ctx *ctx, const float *zyx, int start_y, int end_y);

/* Copies volume rows [start_y,end_y) back into the host volume `zyx` (full
 * D x H x W array; other rows untouched).  Replaces ProjectorGPU.updateFromGPU
 * (ProjectorGPU.java:297-315). */
int tj_get_volume(tj_ctx *ctx, float *zyx, int start_y, int end_y);

/* Sets every voxel to zero (TomoReconstruction2.resetPixels, :414-428). */
int tj_reset_volume(tj_ctx *ctx);

/* VoxelProjector3D.setPositivityConstraint (VoxelProjector3D.java:102-104;
 * ProjectorGPU.setPositivityConstraint, ProjectorGPU.java:149-151): changes the
 * constraint of an open context (doGPURec sets it after initForIterative). */
int tj_set_positivity(tj_ctx *ctx, int on);

/* VoxelProjector3D.createErrorVolume (VoxelProjector3D.java:858-864) when `on`
 * (a zero volume of the reconstruction's size on the device), or drops it.
 * While it exists, every backprojection of differences (tj_backproject after
 * tj_project_and_difference, and every subset of tj_run_ossart) also squares
 * the differences in place (convertDiffToErrorImage, :315-324; NaN pixels stay
 * NaN) and backprojects the squares into it (backproject, :181-195), after the
 * reconstruction's own backprojection of the same differences. */
int tj_set_error_volume(tj_ctx *ctx, int on);

/* Copies the error volume to the host in ImageStack layout (the data
 * VoxelProjector3D.endOfIteration saves as a TIFF stack, :828-836). */
int tj_get_error_volume(tj_ctx *ctx, float *zyx);

/* VoxelProjector3D.clearErrorVolume (:842-849); no-op without an error volume. */
int tj_clear_error_volume(tj_ctx *ctx);

/* Projector.addProjection(int index) (Projector.java:130; VoxelProjector3D.java:111-118):
 * appends resident projection `index` to the ordered list of added projections. */
int tj_add_projection(tj_ctx *ctx, int index);

/* Projector.clearAllProjections() (Projector.java:132; VoxelProjector3D.java:123-129). */
int tj_clear_projections(tj_ctx *ctx);

/* Number of projections currently added. */
int tj_num_added(tj_ctx *ctx);

/* Projector.project(startY,endY) (Projector.java:144; VoxelProjector3D.java:333-376):
 * forward-projects the volume for every added projection; results are read with
 * tj_get_projection. */
int tj_project(tj_ctx *ctx, int start_y, int end_y);

/* Projector.projectAndDifference(factor,startY,endY) (Projector.java:150;
 * VoxelProjector3D.java:387-433): projects every added projection, forms
 * diff = (float)(((exp - th) / norm) * factor) and writes one FP64 error
 * (sum of diff^2, computed before the float cast) per added projection into
 * `errors` (length tj_num_added). */
int tj_project_and_difference(tj_ctx *ctx, double factor, int start_y, int end_y, double *errors);

/* Projector.getProjection(int) (Projector.java:156; VoxelProjector3D.java:131-133):
 * copies the projection (and, if `norm` is not NULL, the norm image) of added
 * slot `slot` from the last tj_project / tj_project_and_difference.  Either
 * output may be NULL. */
int tj_get_projection(tj_ctx *ctx, int slot, float *proj, float *norm);

/* Copies the difference image of added slot `slot` from the last
 * tj_project_and_difference (VoxelProjector3D.getDifferenceImage, :139-141). */
int tj_get_difference(tj_ctx *ctx, int slot, float *diff);

/* Allocation-free projectAndDifference.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * tj_project_and_difference on a context holding this volume and these images
 * added in order 0..n-1: the n difference images bit for bit, the FP64 errors
 * summed in another order (a relative difference of order 1e-15).  Inputs and
 * outputs are the caller's device buffers: `vol_yzx` the volume in the
 * library's resident layout (row y, slice z, x fastest: vol_h x vol_d x vol_w),
 * `measured` and `diff` n x proj_h x proj_w; `euler` (n x 9) and `errors` (n)
 * are host arrays.  Every workspace lives in one device array of
 * TJ_NOALLOC_SCRATCH_BYTES per device, held until the process ends; the call
 * allocates nothing.  Covers the geometry of the LDS forward kernel: one vertical
 * tilt axis, scale 1, whole volume rows, a finite volume and no long-object
 * compensation (otherwise TJ_EUNSUP); a size whose workspaces exceed the array
 * (tj_noalloc_scratch_bytes) gives TJ_ENOMEM.  Runs on `stream` (a hipStream_t;
 * NULL: the null stream) and returns when the work is done.  Calls on one device
 * share the array and are serialised by the library; the array keeps the last
 * call's contents and nothing carries over from one call to the next. */
#define TJ_NOALLOC_SCRATCH_BYTES (16u << 20)
size_t tj_noalloc_scratch_bytes(const tj_geometry *g, int n);
int tj_project_and_difference_noalloc(int device, const tj_geometry *g, const float *vol_yzx, const float *measured,
                                      const double *euler, int n, double factor, int start_y, int end_y, float *diff,
                                      double *errors, void *stream);

/* Projector.backproject(startY,endY) (Projector.java:138; VoxelProjector3D.java:155-308):
 * adds the difference images of every added projection (or, when no difference
 * was computed since the last clear, the projections themselves, weighted if a
 * weighting filter is set) into volume rows [start_y,end_y). */
int tj_backproject(tj_ctx *ctx, int start_y, int end_y);

/* Projector.setWeightingFilter(new FFTWeighting(ts, diameter))
 * (Projector.java:48-50; FFTWeighting.java:21-65): Radermacher exact weighting
 * for weighted backprojection.  tilt_deg / axis_deg: n angles (axis_deg may be
 * NULL for a vertical axis).  diameter = weightingRadius * ts.getWidth()
 * (ReconstructionParameters.java:320).  diameter <= 0 removes the filter. */
int tj_set_weighting(tj_ctx *ctx, double diameter, const double *tilt_deg, const double *axis_deg, int n);

/* Ramp x Hamming filter for weighted backprojection.  No counterpart in TomoJ
 * (which has only the Radermacher weighting); provided for the WBP pass of the
 * production workload.  Filters each detector row in Fourier space with
 * |k|/(N/2) * (0.54 + 0.46 cos(pi k/(N/2)))-style weights (see the library
 * source for the exact table).  cutoff in (0,1] scales the Hamming window. */
int tj_set_ramp_filter(tj_ctx *ctx, int hamming, double cutoff);

/* Copies the current weighting mask (proj_h x proj_w floats). */
int tj_get_weighting_mask(tj_ctx *ctx, float *mask);

/* Applies the current weighting filter to one image in place (host in/out);
 * FFTWeighting.weighting(ImageProcessor) (FFTWeighting.java:87-112). */
int tj_apply_weighting(tj_ctx *ctx, float *image);

/* Projector.startOfIteration / endOfIteration (Projector.java:152-154). */
int tj_start_iteration(tj_ctx *ctx);
int tj_end_iteration(tj_ctx *ctx);

/* FISTA wrapper: ReconstructionParameters.getProjector returns
 * new FistaProjector3D(proj) when fista is set (ReconstructionParameters.java:331).
 * With `on`, tj_start_iteration keeps the volume and tj_end_iteration applies
 * FistaProjector3D.endOfIteration (FistaProjector3D.java:43-67); tj_run_ossart
 * does both around each iteration.  Resets t to 0. */
int tj_set_fista(tj_ctx *ctx, int on);

/* TVM regularisation: the projector AdvancedReconstructionParameters.getProjector
 * builds for the TVM type (TomoJ_AdvancedReconstruction/.../tomography/
 * AdvancedReconstructionParameters.java:179-193) is TVMProjector3D
 * (.../projectors/TVMProjector3D.java:24-157), whose endOfIteration denoises the
 * whole volume with TotalVariationDenoising_3D.denoise3D(stack, theta, g, dt,
 * iterations) (Filters/.../TotalVariationDenoising_3D.java:84-202), dt being
 * (float) regularizationLambda (:152).  With `iterations` > 0, tj_end_iteration
 * and every iteration of tj_run_ossart end with that step (before the FISTA
 * extrapolation when both are set: FistaProjector3D wraps TVMProjector3D);
 * 0 turns it off.  theta and g are cast to float as setTvmTheta / setTvmG do
 * (:64-74).  Holds 16 bytes per voxel of device memory while on. */
int tj_set_tvm(tj_ctx *ctx, int iterations, double theta, double g, double lambda);

/* ReconstructionParameters.getAvailableIndexes (ReconstructionParameters.java:212-256)
 * for `nproj` images: projection_type 11 (EVEN_PROJECTIONS) or 12
 * (ODD_PROJECTIONS) keeps the images of that parity, at most nproj/2 of them;
 * any other type (ALL_PROJECTIONS = 10) keeps all.  Writes the list to `out`
 * (room for nproj) and returns its length, or a negative TJ_E* code. */
int tj_available_indexes(int nproj, int projection_type, int *out);

/* TomoReconstruction2.orderShuffleART (TomoReconstruction2.java:735-781): the
 * order in which an iteration walks `n` images, each next one the remaining
 * image whose third Euler row has the smallest |dot product| with the
 * previous one's (euler: n x 9, row-major).  TomoJ picks the first image with
 * Math.random() (:755); here the caller passes it.  Writes n indexes. */
void tj_order_shuffle(const double *euler, int n, int first, int *order);

/* The images tj_run_ossart and tj_run_wbp walk: the list getAvailableIndexes
 * returns (or setAvailableIndexes sets, :258-261), as TomoReconstruction2.WBP
 * (:500) and orderShuffleART (:736) use it; n = 0 restores all.  With a list,
 * tj_run_ossart keeps TomoJ's arithmetic on the full count N: the order
 * chain uses the Euler matrices of images 0..n-1 (:741) and emits the listed
 * images, update 0 means N, a subset is flushed when it holds `update` images
 * or at position N-1 (:662), so a last partial subset is never projected and
 * with update 0 nothing is, and err[0] divides by N (:726).  first_index is a
 * position in the list.  Checked against the loaded projections at run time. */
int tj_set_available_indexes(tj_ctx *ctx, const int *indexes, int n);

/* Fourier shell correlation of two reconstructions:
 * ResolutionEstimation.fsc(recFSCeven, recFSCodd)
 * (tomography/ResolutionEstimation.java:456-557) of the resident volumes of
 * `even` and `odd` (same size, same device).  out = fsc[4][W/2+1] row by row:
 * frequency i/W, FSC, the 2/sqrt(n) threshold curve, mean |E - O| per shell.
 * Both volumes are multiplied by TomoJ's soft spherical mask (radius = the
 * slice count), transformed in 3D in FP32 and the shells summed in FP64 (the
 * atomics make the last bits of the sums order-dependent).  Needs 24 bytes
 * per voxel of scratch device memory during the call; W/2+1 <= 1638 shells,
 * H and D <= 65535. */
int tj_fsc(tj_ctx *even, tj_ctx *odd, double *out);

/* Spectral SNR of a reconstruction: ResolutionEstimation.SSNR(false)
 * (tomography/ResolutionEstimation.java:167-260 with addToVectors1D :581-697).
 * `signal` holds the resident (supplied) tilt series and its reconstruction,
 * `noise` a tilt series of Gaussian noise (generateSSNRNoiseData, :93-112)
 * reconstructed with the same parameters; same sizes.  For every tilt both
 * reconstructions are re-projected (ReconstructionThreadManager.getProjection,
 * :257-270), NaN set to 0, and the power spectra of measured - projected and
 * of projected summed over rings ring_width wide (TomoJ: 4).  out (9 x
 * (W/2+1) doubles, row by row) receives TomoJ's nine rows: index, frequency /
 * pixel_size, corrected SSNR in dB (-1000 where SSNR <= 1), signal SNR, signal
 * and noise powers in dB, noise SNR and its powers in dB; *nfreq = the number
 * of columns filled (W/2+1).  FP32 transforms (TomoJ's are FP64), FP64 sums.
 * vssnr (may be NULL): n x proj_h x proj_w floats, the per-pixel SSNR maps of
 * the computeVSSNR path (:648-663), 10 log10(max(ISSNR/alpha - 1, 0) + 1) in
 * swapQuadrant2D's order (:720-733; an odd width's last column stays 0): the
 * tilt series SSNR(true) reconstructs into the VSSNR volume. */
int tj_ssnr(tj_ctx *signal, tj_ctx *noise, int ring_width, double pixel_size, double *out, int *nfreq,
            float *vssnr);

/* The same for two host volumes in ImageStack layout [D][H][W] (the static
 * fsc(ImagePlus, ImagePlus)); copies them to `device` for the call. */
int tj_fsc_volumes(int device, const float *even_zyx, const float *odd_zyx, int W, int H, int D, double *out);

/* Bayesian regularisation: the projector getProjector builds for the BAYESIAN
 * type (AdvancedReconstructionParameters.java:194-203) is
 * BayesianVoxelProjector3D (TomoJ_AdvancedReconstruction/.../projectors/
 * BayesianVoxelProjector3D.java:25-204).  With `on`: tj_start_iteration (and
 * each iteration of tj_run_ossart) sets a device distance volume to
 * -lambda x the volume convolved with the [1,-2,1]^3 stencil, zero outside
 * (startOfIteration / convolve, :134-198); every backprojection of
 * differences adds into the distance volume instead of the reconstruction,
 * and a backprojection of the original images (no differences, tj_run_wbp)
 * adds nothing (:75-110); tj_end_iteration multiplies the distance volume by
 * (float) alpha, adds it to the volume and applies the positivity constraint
 * when the geometry has it (:121-132).  alpha = regularizationAlpha (the
 * relaxation coefficient), lambda = regularizationWeight.  TomoJ never wraps
 * this projector in FISTA; the hooks still compose if asked.  Refused while
 * the TVM hook is on (and tj_set_tvm is refused while this one is).  Holds 4
 * bytes per voxel of device memory while on. */
int tj_set_bayesian(tj_ctx *ctx, int on, double alpha, double lambda);

/* BgART: the projector getProjector builds for the BGART type
 * (AdvancedReconstructionParameters.java:145-155, 214-218) is
 * BgARTVoxelProjector3D (TomoJ_AdvancedReconstruction/.../projectors/
 * BgARTVoxelProjector3D.java:13-152).  With `on`, the collapse filter runs
 * once at once (the constructor, :20-27) and at every tj_end_iteration
 * (:119-121): the stack's min, max and 256-bin histogram (ImageJ
 * StackStatistics), the background mv = min + (first fullest bin) x binSize,
 * or bg_value when it is not NaN, sigma from the bins below it, and every
 * voxel below mv + k sigma (dark_bg) or above mv - k sigma (otherwise) set to
 * (float) mv (collapseFilter, :30-78).  With `median`, tj_start_iteration (and
 * each iteration of tj_run_ossart) first replaces every voxel with the median
 * of the cross it forms with its neighbours -+1 in z (0 beyond the volume) and
 * -+1, -+W in its slice image's row-major index, those wrapping across rows
 * and read as 0 where TomoJ's tests idx+1 < W*H, idx-1 > 0, idx-W > 0,
 * idx+W < W*H fail (startOfIteration, :123-151).  Refused while TVM or
 * Bayesian is on.  Holds 4 bytes per voxel of device memory while on. */
int tj_set_bgart(tj_ctx *ctx, int on, double k, int dark_bg, int median, double bg_value);

/* TotalVariationDenoising_3D.denoise3D(stack, theta, g, dt, iterations)
 * (TotalVariationDenoising_3D.java:84-202) of the resident volume, once: the
 * plugin's run (:66-73) with the volume as the stack. */
int tj_tv_denoise(tj_ctx *ctx, float theta, float g, float dt, int iterations);

/* Projector.getCompletion (Projector.java:122-124); may be called from any thread. */
double tj_get_completion(tj_ctx *ctx);

/* TomoReconstruction2.interrupt() (:137-140): the running tj_run_* loop stops at
 * its next subset boundary and returns TJ_ECANCELLED.  Any thread. */
int tj_interrupt(tj_ctx *ctx);

/* TomoReconstruction2.OSSART(ts, projector, params, startY, endY)
 * (TomoReconstruction2.java:548-733) run inside the library over the resident
 * projections: `iterations` iterations of osartIteration with relaxation
 * `lambda` and `update` projections per subset (0 = all: SIRT; 1 = ART).  The
 * projection order of each iteration is orderShuffleART (:735-781) started at
 * `first_index` instead of Math.random().  err (iterations x 5, may be NULL)
 * receives err1 = [avg, min, image of min, max, image of max] per iteration. */
int tj_run_ossart(tj_ctx *ctx, int iterations, double lambda, int update, int first_index,
                  int start_y, int end_y, double *err);

/* TomoReconstruction2.WBP(ts, projector, params, startY, endY)
 * (TomoReconstruction2.java:499-538) over the resident projections: each one
 * weighted with the current filter (if any) and backprojected into the
 * volume (which is not cleared first). */
int tj_run_wbp(tj_ctx *ctx, int start_y, int end_y);

/* Residual norms of the last tj_project_and_difference in FP64: sum over added
 * projections of sum diff^2 (what osartIteration accumulates in `som`). */
int tj_last_error_sum(tj_ctx *ctx, double *som);

/* ---------------------------------------------------------------------------
 * Synthetic inputs (no TomoJ counterpart: TomoJ reads tilt series from disk).
 * Seeded ellipsoid phantom with analytic FP64 line integrals in TomoJ's
 * geometry plus Gaussian noise, generated on the device straight into the
 * resident projection stack.
 * ------------------------------------------------------------------------- */

/* Ellipsoid parameters (k x 16 doubles: density, centre xyz, semi-axes xyz,
 * rotation 3x3 row-major) from a SplitMix64 stream seeded with `seed`, for a
 * W x H x D volume. Host-only; no context needed. */
void tj_phantom_ellipsoids(uint64_t seed, int k, int vol_w, int vol_h, int vol_d, double *out);

/* Generates n analytic projections of the ellipsoids `ell` (k x 16) for Euler
 * matrices `euler` (n x 9), adds N(0, noise_rel * max) noise drawn from
 * `noise_seed`, and makes them the resident projections (as tj_load_projections). */
int tj_synth_projections(tj_ctx *ctx, const double *ell, int k, const double *euler, int n,
                         double noise_rel, uint64_t noise_seed);

/* Projection supply with the identity alignment: what TiltSeries.getPixels(i)
 * returns with normalisation on (TomoJ's default, TiltSeries.java:128),
 * AffineAlignment.applyTransformOnImage (AffineAlignment.java:136-156):
 * every resident projection i becomes (float)((x - mean_i) / stdDev_i) with the
 * statistics of its ROI (TiltSeries.getImageStatistics :809-835, columns
 * round(W cos(tilt_deg[i])) wide, at least min_size).  stats (may be NULL)
 * receives n x {mean, stdDev}. */
int tj_normalize_projections(tj_ctx *ctx, const double *tilt_deg, int min_size, double *stats);

/* Fill of aligned pixels that map outside the original (TiltSeries.FILL_*,
 * TiltSeries.java:48-56). */
#define TJ_FILL_NONE 0
#define TJ_FILL_AVG 1
#define TJ_FILL_NAN 2

/* Projection supply with an affine alignment: what TiltSeries.getPixels(i)
 * returns through AffineAlignment.applyTransformOnImage (AffineAlignment.java:
 * 136-254) for every resident projection, in place.  transforms: n x 6, the
 * forward transform T_i of AffineAlignment.getTransform(i) in
 * java.awt.geom.AffineTransform.getMatrix order (m00, m10, m01, m11, m02, m12);
 * an identity T_i takes the identity branch.  Otherwise pixel (i, j) samples the
 * original bilinearly at T_i^-1 (i - W/2, j - H/2) + (W/2, H/2) (the last column
 * and row nearest), is normalised with the image's statistics when `normalize`,
 * and pixels mapping outside take `fill`.  The statistics are those of
 * TiltSeries.getImageStatistics(i) (:809-835): pass them in stats_in (n x
 * {mean, stdDev}, e.g. from ImageJ), or NULL to compute them on the device over
 * the ROI rectangle's corners mapped by T_i and cast to int: the rectangle
 * they span for transforms without rotation or shear, else the pixels whose
 * centres lie in their polygon (even-odd test; ImageJ's PolygonRoi mask rule,
 * restated).  tilt_deg is needed only then.
 * stats_out (may be NULL) receives the statistics used.  TJ_EINVAL for a
 * non-invertible transform. */
int tj_align_projections(tj_ctx *ctx, const double *transforms, const double *tilt_deg, int min_size,
                         int normalize, int fill, const double *stats_in, double *stats_out);

/* The optional threshold of the projection supply: TiltSeries.getPixels applies
 * applyThreshold (TiltSeries.java:682-684, :720-806) after the alignment when
 * setThreshold(threshold, true) or setThresholdHisteresis(threshold, hysteresis,
 * true) was called (:407-416).  On every resident projection, in place: the
 * levels at the `threshold` and `hysteresis` fractions of ImageJ's 256-bin
 * histogram of the image, values below the first take the fill value
 * (TJ_FILL_NONE 0, TJ_FILL_AVG the image mean, TJ_FILL_NAN), values between the
 * two take it when connected to such a value through index -+1 or -+width.
 * setThreshold leaves the hysteresis at Double.MIN_VALUE (4.9e-324). */
int tj_threshold_projections(tj_ctx *ctx, double threshold, double hysteresis, int fill);

/* ---------------------------------------------------------------------------
 * MRC files (InputOutput/MRC_Reader.java, MRC_Writer.java): the container
 * TomoJ reads tilt series from and writes reconstructions to.  Host only.
 * ------------------------------------------------------------------------- */
typedef struct tj_mrc_info {
    int nx, ny, nz;    /* columns, rows, sections (images)                      */
    int mode;          /* 0 uint8, 1 int16, 2 float32, 6 uint16                  */
    int little_endian; /* byte order found by MRC_Reader's rule (:176-215)       */
    int extra_bytes;   /* extended header size (word 23)                         */
    int fei;           /* extended header of 131072 bytes: FEI tilt angles       */
    float amin, amax, amean; /* header statistics (words 19-21)                  */
} tj_mrc_info;

/* Header of an MRC file, as MRC_Reader.getMRCFileInfo reads it
 * (MRC_Reader.java:176-233).  TJ_EUNSUP for other data modes, TJ_EIO when the
 * file cannot be read. */
int tj_mrc_read_info(const char *path, tj_mrc_info *info);

/* The data as float, nz x ny x nx (ImageStack layout); tilts (may be NULL)
 * receives nz FEI tilt angles (MRC_Reader.readTiltAngles :235-256), or NaN
 * when the file has no FEI extended header. */
int tj_mrc_read(const char *path, float *out, float *tilts);

/* A float volume or stack nz x ny x nx as MRC_Writer writes a float image
 * (MRC_Writer.java:60-160): little-endian, mode 2, header statistics. */
int tj_mrc_write(const char *path, const float *data, int nx, int ny, int nz);

/* Copies resident projection `index` to the host. */
int tj_get_loaded_projection(tj_ctx *ctx, int index, float *out);

/* Timing of the most recent tj_run_ossart / tj_run_wbp, from HIP events on the
 * context's stream: total milliseconds, and the summed milliseconds and launch
 * counts of the forward-projection kernel (in the fused loop: the forward kernel
 * that writes the differences), the residual kernels (difference images and
 * error sums; in the fused loop the fold of the error partials), the
 * backprojection's preparation kernels (row table and, for the LDS kernel, the
 * FP64 pair images) and the backprojection kernel itself. */
typedef struct {
    double total_ms;
    double fp_ms, res_ms, bp_ms, other_ms;
    int fp_launches, res_launches, bp_launches;
    double bpprep_ms;
    int bpprep_launches;
    int fused_subsets;  /* subsets whose projectAndDifference ran as the fused
                           forward kernel (counted with timing on or off) */
    int fast_bp_launches; /* backprojections the fast form took (timing on) */
} tj_timing;
int tj_get_timing(tj_ctx *ctx, tj_timing *t);

/* Turns per-kernel event timing on/off for tj_run_* (off by default). */
int tj_enable_timing(tj_ctx *ctx, int on);

/* Kernel selection and launch shapes (no TomoJ counterpart: TomoJ lets the
 * OpenCL driver pick the work-group, GPUDevice.java:303-313).  Used by the
 * autotuner; every choice gives bit-identical results.  kernel: 0 automatic,
 * 1 any geometry, 2 single tilt axis, 3 single tilt axis with LDS staging
 * (used only where its preconditions hold; otherwise 2). */
typedef struct {
    int fp_kernel;
    int fp_rows;    /* detector rows per thread in kernel 2: 1, 2, 4, 8 (kernel 3: 4) */
    int fp_rays;    /* projections per thread in kernel 3: 2, 4, 8                 */
    int fp_zchunk;  /* slices per LDS chunk in kernel 3 (0: as many as fit)        */
    int bp_kernel;
    int bp_rows;    /* volume rows per thread: 2, 4 (1, 8 for kernel 2; 8 for 3)   */
    int bp_zchunk;  /* slices per thread: 4, 8 (16 with 2 rows in kernel 3)        */
    int bp_group;   /* images per LDS pass in kernel 3: 2, 3, 4                    */
    int bp_fast;    /* kernel 3's fast form for a finite volume without positivity
                       when every image row it meets is finite at an integer
                       position: 0 off, 1..7 its shape (rows x slices x images:
                       1 4x8x2, 2 4x8x1, 3 8x8x1, 4 4x16x1, 5 2x16x2, 6 8x4x2,
                       7 4x4x2)                                                  */
    int fp_groups;  /* kernel 3: groups of fp_rays projections per block, 1, 2 or
                       4 (one 256-thread part each, sharing one staged slab);
                       0 automatic: 4 with the double-buffered slab, else 4 for
                       launches of up to 4 groups and 2 for longer ones       */
    int fp_prefetch; /* kernel 3's slab staging: 0 double-buffered where it fits
                       (the next chunk's loads in flight while the current one
                       is projected), 1 a single buffer                        */
} tj_tuning;
int tj_set_tuning(tj_ctx *ctx, const tj_tuning *t);
int tj_get_tuning(tj_ctx *ctx, tj_tuning *t);

/* ---------------------------------------------------------------------------
 * Job engine: precision, batched and serving forms.  No TomoJ counterpart
 * (TomoJ reconstructs one tilt series per ReconstructionThreadManager).  A
 * tj_batch holds any number of independent SIRT jobs resident on one device
 * within a fixed HBM budget; tj_batch_run advances a set of them together:
 * per iteration one persistent forward-projection launch per storage type
 * (with the residual fused: TomoJ's projectAndDifference,
 * VoxelProjector3D.java:387-433), one FP64 error-reduction launch and one
 * persistent backprojection launch per storage type and positivity setting
 * (VoxelProjector3D.java:155-308), each walking a table of tiles over all the
 * jobs.  Jobs are tilt series about the vertical axis with TomoJ's default
 * centres (TiltSeries.java:161-170); SIRT is osartIteration with update = N
 * (TomoReconstruction2.java:636-733, factor (float)(lambda/N)).  Positions are
 * closed-form in the arithmetic type instead of Java's incremental doubles, so
 * results match the FP64 oracle to a stated tolerance rather than bit for bit.
 * A job's result does not depend on which other jobs share its launches.
 * Failures are reported like every other call (status code, tj_last_error).
 * ------------------------------------------------------------------------- */
typedef struct tj_batch tj_batch;

typedef struct {
    int vol_w, vol_h, vol_d;   /* reconstruction size (width <= 4096)              */
    int proj_w, proj_h;        /* detector size; proj_h <= vol_h, difference even   */
    int n_proj;                /* projections                                       */
    int storage;               /* TJ_STORAGE_*                                      */
    int positivity;            /* clamp at 0 after each image (VoxelProjector3D:299) */
    double relaxation;         /* lambda                                            */
    int max_iterations;        /* length of the error history kept on the device    */
    int fp32_master;           /* BF16/FP16 storage: keep the SIRT iterate in an FP32
                                  master volume; the 16-bit copy feeds the forward
                                  projection (half its traffic), the backprojection
                                  updates both.  Ignored for FP32/FP64 storage.      */
    int fp32_projections;      /* BF16/FP16 storage: keep the measured projections in
                                  FP32 (they are read once per iteration).           */
    int reserved;
} tj_job_desc;

/* Tile and slab shapes of the job kernels (picked by tj_batch_autotune). */
typedef struct {
    int fp_rays;        /* projections per forward tile: 2 or 4                     */
    int fp_lds_kb;      /* LDS cap per forward block in KiB (16..152): slices per
                           slab; a launch takes less where more would cost
                           resident blocks (the registers' limit)               */
    int bp_rows;        /* volume rows per back tile: 2, 4 or 8                      */
    int bp_zchunk;      /* slices per back tile: 4, 8 or 16                          */
    int bp_group;       /* images per LDS pass (derived from the shape)              */
    int blocks_per_cu;  /* persistent blocks per compute unit: 1..16, capped at
                           what one CU holds of each kernel                      */
    int fp_packed;      /* BF16/FP16 storage: forward projection on packed 16-bit
                           pairs with v_dot2 (8 rows per tile) instead of FP32;
                           2: the same on a double-buffered slab shared by four
                           projection groups per 1024-thread block (even widths
                           up to 2048; other launches take form 1)            */
    int fp_cols;        /* packed forward: 256-column tiles per block, 1 or 2 (two
                           share one staged slab)                                 */
    int fp_shared;      /* FP32-arithmetic forward (FP32 storage, or 16-bit without
                           the packed form): 1 (default) four projection groups
                           of 4 rays per 1024-thread block sharing a
                           double-buffered slab (widths up to 3071; wider
                           launches take form 0), 0 one group per 256-thread
                           block                                                  */
} tj_batch_tuning;

#define TJ_RUN_GRAPH 1  /* tj_batch_run: capture the iterations in one HIP graph */

/* Device bytes a job of this description takes (0 if the description is invalid). */
size_t tj_job_bytes(const tj_job_desc *d);
int tj_batch_open(int device, size_t hbm_budget_bytes, tj_batch **out);
int tj_batch_close(tj_batch *b);
size_t tj_batch_used(const tj_batch *b);
size_t tj_batch_budget(const tj_batch *b);
/* Admits a job (TJ_ENOMEM when it does not fit the budget); euler is n_proj x 9
 * (AffineAlignment.getEulerMatrix per tilt).  Volume and projections start at 0. */
int tj_batch_add_job(tj_batch *b, const tj_job_desc *d, const double *euler, int *job);
int tj_batch_remove_job(tj_batch *b, int job);
/* Measured projections float[n][ph][pw], converted to the job's storage. */
int tj_batch_set_projections(tj_batch *b, int job, const float *stack);
/* Analytic projections of an ellipsoid phantom plus noise (as tj_synth_projections). */
int tj_batch_synth_job(tj_batch *b, int job, const double *ell, int k, double noise_rel, uint64_t seed);
int tj_batch_get_projection(tj_batch *b, int job, int index, float *out);
/* Difference image `index` of the job's last forward pass, widened to double. */
int tj_batch_get_difference(tj_batch *b, int job, int index, double *out);
/* Volume in TomoJ's ImageStack layout float[D][H][W]. */
int tj_batch_set_volume(tj_batch *b, int job, const float *zyx);
int tj_batch_get_volume(tj_batch *b, int job, float *zyx);
int tj_batch_get_volume_f64(tj_batch *b, int job, double *zyx);  /* FP64 jobs only */
int tj_batch_reset_volume(tj_batch *b, int job);
/* Changes a job's storage type in place (e.g. BF16 -> FP32 before refinement);
 * the volume is taken from the FP32 master when there is one.  Projections
 * are converted from their stored values: supply the FP32 originals again
 * (tj_batch_set_projections / tj_batch_synth_job) to refine against them. */
int tj_batch_convert(tj_batch *b, int job, int storage);
/* `iterations` SIRT iterations of every listed job; synchronous. */
int tj_batch_run(tj_batch *b, const int *jobs, int njobs, int iterations, int flags);
double tj_batch_last_run_ms(const tj_batch *b);
/* Returns the number of iterations the job has done and copies the sums of
 * squared differences of the last min(that, cap, max_iterations) of them. */
int tj_batch_errors(tj_batch *b, int job, double *err, int cap);
int tj_batch_set_tuning(tj_batch *b, const tj_batch_tuning *t);
int tj_batch_get_tuning(const tj_batch *b, tj_batch_tuning *t);
/* Times candidate tile/slab shapes on the listed jobs (volumes are restored)
 * and keeps the fastest; returns it and its forward + back milliseconds. */
int tj_batch_autotune(tj_batch *b, const int *jobs, int njobs, tj_batch_tuning *best, double *best_ms);

#ifdef __cplusplus
}
#endif

#endif /* TOMOJ_HIP_H */
