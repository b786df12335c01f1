// tj_device.h -- structures shared by the host side (tj_api.hip) and the kernels
// (tj_kernels.hip).  Device layout of the resident data:
//   volume       [y][z][x]   (one contiguous XZ plane per volume row y, so a
//                             Y-slab [y0,y1) is one contiguous block)
//   projections  [n][j][i]   (measured images, TomoJ's [index][row][column])
//   slot images  [slot][j][i] (forward projections, norms, differences of the
//                             projections added since the last clear)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Geometry in the form the kernels use (values taken from tj_geometry).
struct TjGeomDev {
    int vw, vh, vd;       // volume width (x), height (y), depth (z)
    int pw, ph;           // detector width, height
    int positivity;
    int pad0;
    double pc0, pc1;      // projectionCenters[0..1]: forward projection
    double pc2, pc3;      // projectionCenters[2..3]: backprojection
    double vcx, vcy, vcz; // volumeCenters
    double scale;
    double bxmax, bymax;  // forward-projection borders: volumeWidth + ACCURACY evaluated
                          // in float (VoxelProjector3D.java:700-701), i.e. exactly W, H
};

// Per resident projection: Euler matrix plus the constants projectJosephPartial
// derives from it (VoxelProjector3D.java:682-694), computed on the host in the
// same operation order as the Java code.
struct TjProj {
    double e[9];
    double icf, zcf, divy, yfx, zfx, divx, dirx, diry, norminc;
    int axis;   // 1: the ray keeps a constant y and x does not depend on y
                // (e01 == e10 == e12 == 0), which the fast kernels exploit
    int pad;
};

// Per (slot, volume row) backprojection row table for the axis-aligned kernel:
// yy = (y*e11)*scale + ciy is the same for every voxel of the row.
struct TjBpRow {
    double yf;     // yy - (int)yy
    int yb, y1;    // row and clamped next row of the image
    int valid;     // 0 <= yy < ph
    int shortcut;  // the y-interpolation reduces to the lower row exactly
    int finite;    // the lower image row holds only finite values
    int pad;
};

// ACCURACY (VoxelProjector3D.java:33) is a float constant widened to double.
#define TJ_ACCURACY ((double)1e-10f)

#define TJ_BLOCK 256
#define TJ_SU 4  // columns per lane per LDS staging batch (loads in flight)

// Stores the message tj_last_error() returns on the calling thread.
void tj_record_error(const char* msg);

// TomoReconstruction2.orderShuffleART with the first image fixed (host code;
// part of the C ABI, declared in tomoj_hip.h as well).
extern "C" void tj_order_shuffle(const double* euler, int n, int first, int* order);

// Host helpers shared by the context API (tj_api.hip) and the allocation-free
// entry point (tj_noalloc.hip); tj_geom_to_dev and tj_fp_start_row, which take
// the ABI's tj_geometry, are declared where tomoj_hip.h is included.
TjProj tj_make_proj(const double* r);
bool tj_fp_lds_ok(const TjGeomDev& g, const TjProj* hpp, const int* idx, int n, int start_row);
