// Kernel argument blocks and launchers of the MI355X PSTD core.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pstd.h"

namespace pstd {

// Per-axis segment table, passed by value (PSTD_QScat.cpp:277-331 ranges).
struct SegTable {
    int n;
    int fft1[PSTD_MAX_SEG];
    int core1[PSTD_MAX_SEG];
    int ncore[PSTD_MAX_SEG];
    int mb[PSTD_MAX_SEG];       // 16-row blocks actually used by segment
    int opoff[PSTD_MAX_SEG];    // operator fragment offset (elements)
};

// x and y passes: rows l of a segment are planes/rows with a common
// contiguous column index.  psi0[core rows] += -i * (A psi_seg).
template <typename T>
struct PlaneArgs {
    const T *psi;
    T *psi0;
    const T *op;
    long row_stride;     // complex elements between consecutive rows l
    long outer_stride;   // complex elements between outer slabs
    int outer_count;     // 1 (x pass) or Nx (y pass)
    int ncols;           // complex points per row
    int col_tiles;
    int L, KB, MB;
    SegTable seg;
};

// z pass fused with the pointwise terms of Update (PSTD_QScat.cpp:1268-1435).
template <typename T>
struct LineArgs {
    const T *psi;
    T *psi0;
    const T *op;
    long nlines;         // Nx * Ny
    int Ny, Nz, L, KB, MB, line_tiles;
    SegTable seg;
    // potential box
    const T *vbox;
    int vb1[3], vb2[3];
    T vfac;              // ramp factor * (V~ already holds 2 dtau / E0)
    int use_v;
    // TF/SF source tables (FP64)
    const double *tfP[3];
    const double2 *tfA[3];
    const double2 *tfE[3];
    int k1a[3], k1b[3], k2a[3], k2b[3];
    double sin_t, cos_t;
    // ABC
    const T *g[3];
};

struct SurfFace {
    long stride_l, stride_b, stride_c;   // complex elements
    long fft1;                           // global start of the owning segment
    int li, nb, nc;
    const double *ps, *fd;               // length L
    double2 *acc;                        // [3][nb*nc]: psi, PS, FD
};
struct SurfArgs {          // the two faces of one axis
    SurfFace f[2];
    int L;
    double ph_r, ph_i;                   // e^{i tau}
};

// launchers (pstd_kernels.hip)
template <typename T> hipError_t launch_plane(const PlaneArgs<T> &a, int grid_x, hipStream_t s);
template <typename T> hipError_t launch_line(const LineArgs<T> &a, int grid_x, hipStream_t s);
template <typename T> hipError_t launch_surface(const T *psi, const SurfArgs &a, long max_pts, hipStream_t s);
template <typename T> size_t pass_lds_bytes(int MB, int KB);
template <typename T> int pass_occupancy(int MB, int KB, bool line);
hipError_t configure_kernels(int max_lds);

}  // namespace pstd
