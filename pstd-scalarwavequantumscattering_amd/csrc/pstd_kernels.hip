// HIP kernels of the PSTD_QScat time step for MI355X (gfx950).
//
// The reference computes each axis term of Update (PSTD_QScat.cpp:1448-1537)
// as one MKL FFT pair per local line.  Here the local-Fourier second
// derivative of a windowed segment is a real L x L circulant (D2 =
// IFFT diag(kappa^2) FFT), so every axis pass is a batched real GEMM
//     U[m][col] = sum_l A[m][l] psi[fft1 + l][col],  A = alpha D2[core,:] W
// with Re and Im as separate columns, run on MFMA (16x16x4, FP64 or FP32).
// The overlap exchange (PSTD_QScat.cpp:1539-1731) disappears: a segment's
// halo is W(l) times the neighbour's core, i.e. the global array, and W is
// folded into A.
#include "pstd_kernels.hpp"

namespace pstd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA.  Lane l supplies A[l%16][l/16] and B[l/16][l%16].  Element q
// of the result holds D[row(l/16, q)][l%16] with row = 4*(l/16) + q for the
// FP32 instruction and row = 4*q + l/16 for the FP64 one (both measured on
// gfx950 with an in-bounds probe).
template <typename T> struct VT;
template <> struct VT<double> {
    typedef double2 c2;
    typedef v4d acc;
    static __device__ __forceinline__ int row(int g, int q) { return 4 * q + g; }
};
template <> struct VT<float> {
    typedef float2 c2;
    typedef v4f acc;
    static __device__ __forceinline__ int row(int g, int q) { return 4 * g + q; }
};

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int PT = 32;   // complex points per x/y tile (two 16-point blocks)
constexpr int LT = 32;   // lines per z tile (two 16-line blocks)

__host__ __device__ constexpr size_t align16(size_t n) { return (n + 15) & ~size_t(15); }

template <typename T> size_t plane_lds_bytes(int MB, int KB) {
    return align16((size_t)MB * KB * 64 * sizeof(T)) + (size_t)4 * KB * PT * 2 * sizeof(T);
}
template <typename T> size_t line_lds_bytes(int MB, int KB) {
    return align16((size_t)MB * KB * 64 * sizeof(T)) + (size_t)LT * (4 * KB + 1) * 2 * sizeof(T);
}

// ---------------------------------------------------------------------------
// x / y pass.  Block = one segment (blockIdx.y) and a strided set of column
// tiles; wave w owns 16-row block w of the segment's core.  The operator
// fragments stay in LDS for the life of the block.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_lap_plane(PlaneArgs<T> a) {
    typedef typename VT<T>::c2 C2;
    typedef typename VT<T>::acc ACC;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int r = blockIdx.y;
    const int KB = a.KB, Kp = 4 * KB;
    T *As = reinterpret_cast<T *>(smem);
    C2 *Bs = reinterpret_cast<C2 *>(smem + align16((size_t)a.MB * KB * 64 * sizeof(T)));
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mbr = a.seg.mb[r];
    const T *op = a.op + a.seg.opoff[r];
    for (int i = tid; i < mbr * KB * 64; i += nthr) As[i] = op[i];
    const long fft1 = a.seg.fft1[r], core1 = a.seg.core1[r];
    const int ncore = a.seg.ncore[r];
    const int ntiles = a.outer_count * a.col_tiles;
    const C2 *psi = reinterpret_cast<const C2 *>(a.psi);
    C2 *psi0 = reinterpret_cast<C2 *>(a.psi0);
    const int kq = lane >> 4, cq = lane & 15;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int o = t / a.col_tiles;
        const long c0 = (long)(t - o * a.col_tiles) * PT;
        const C2 *src = psi + (long)o * a.outer_stride + fft1 * a.row_stride + c0;
        __syncthreads();
        for (int i = tid; i < Kp * PT; i += nthr) {
            const int l = i / PT, c = i - l * PT;
            C2 v;
            v.x = 0;
            v.y = 0;
            if (l < a.L && c0 + c < a.ncols) v = src[(long)l * a.row_stride + c];
            Bs[i] = v;
        }
        __syncthreads();
        if (wave < mbr) {
            ACC ar0 = {0, 0, 0, 0}, ai0 = {0, 0, 0, 0}, ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
            const T *Aw = As + wave * KB * 64 + lane;
            const C2 *Bw = Bs + kq * PT + cq;
#pragma unroll 2
            for (int kb = 0; kb < KB; ++kb) {
                const T af = Aw[kb * 64];
                const C2 b0 = Bw[kb * 4 * PT];
                const C2 b1 = Bw[kb * 4 * PT + 16];
                ar0 = mfma16(af, b0.x, ar0);
                ai0 = mfma16(af, b0.y, ai0);
                ar1 = mfma16(af, b1.x, ar1);
                ai1 = mfma16(af, b1.y, ai1);
            }
            // psi0 += -i U  (alpha is folded into the operator)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = wave * 16 + VT<T>::row(kq, q);
                if (m >= ncore) continue;
                C2 *dst = psi0 + (long)o * a.outer_stride + (core1 + m) * a.row_stride + c0;
                if (c0 + cq < a.ncols) {
                    C2 v = dst[cq];
                    v.x += ai0[q];
                    v.y -= ar0[q];
                    dst[cq] = v;
                }
                if (c0 + 16 + cq < a.ncols) {
                    C2 v = dst[16 + cq];
                    v.x += ai1[q];
                    v.y -= ar1[q];
                    dst[16 + cq] = v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// z pass (contiguous lines) with the rest of Update fused into the
// epilogue: potential with its ramp (PSTD_QScat.cpp:1268-1303), TF/SF source
// (:1306-1419, regenerated from separable FP64 tables and the FP64 time
// phase) and the ABC mask (:1421-1435).  The MFMA roles are swapped so that
// each lane ends with both components of one output point and 16 lanes
// cover 16 consecutive z.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_lap_line(LineArgs<T> a) {
    typedef typename VT<T>::c2 C2;
    typedef typename VT<T>::acc ACC;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int r = blockIdx.y;
    const int KB = a.KB, Kp = 4 * KB, LS = Kp + 1;
    T *As = reinterpret_cast<T *>(smem);
    C2 *Ls = reinterpret_cast<C2 *>(smem + align16((size_t)a.MB * KB * 64 * sizeof(T)));
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mbr = a.seg.mb[r];
    const T *op = a.op + a.seg.opoff[r];
    for (int i = tid; i < mbr * KB * 64; i += nthr) As[i] = op[i];
    const long fft1 = a.seg.fft1[r];
    const int core1 = a.seg.core1[r], ncore = a.seg.ncore[r];
    const int i1 = core1 - (int)fft1;
    const C2 *psi = reinterpret_cast<const C2 *>(a.psi);
    C2 *psi0 = reinterpret_cast<C2 *>(a.psi0);
    const int kq = lane >> 4, cq = lane & 15;
    for (int t = blockIdx.x; t < a.line_tiles; t += gridDim.x) {
        const long q0 = (long)t * LT;
        __syncthreads();
        for (int i = tid; i < LT * Kp; i += nthr) {
            const int line = i / Kp, l = i - line * Kp;
            C2 v;
            v.x = 0;
            v.y = 0;
            if (l < a.L && q0 + line < a.nlines) v = psi[(q0 + line) * a.Nz + fft1 + l];
            Ls[line * LS + l] = v;
        }
        __syncthreads();
        if (wave >= mbr) continue;
        ACC ar0 = {0, 0, 0, 0}, ai0 = {0, 0, 0, 0}, ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
        const T *Aw = As + wave * KB * 64 + lane;
        const C2 *L0 = Ls + cq * LS + kq;
        const C2 *L1 = Ls + (16 + cq) * LS + kq;
#pragma unroll 2
        for (int kb = 0; kb < KB; ++kb) {
            const T af = Aw[kb * 64];
            const C2 p0 = L0[4 * kb];
            const C2 p1 = L1[4 * kb];
            ar0 = mfma16(p0.x, af, ar0);
            ai0 = mfma16(p0.y, af, ai0);
            ar1 = mfma16(p1.x, af, ar1);
            ai1 = mfma16(p1.y, af, ai1);
        }
        const int m = wave * 16 + cq;
        if (m >= ncore) continue;
        const int z = core1 + m;
        const T gz = a.g[2][z];
        const bool zin_v = a.use_v && z >= a.vb1[2] && z <= a.vb2[2];
        const bool zin_box = z >= a.k1a[2] && z <= a.k2b[2];
        const bool zin_core = z > a.k1b[2] && z < a.k2a[2];
#pragma unroll
        for (int hb = 0; hb < 2; ++hb) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int line = hb * 16 + VT<T>::row(kq, q);
                const long qg = q0 + line;
                if (qg >= a.nlines) continue;
                const int x = (int)(qg / a.Ny), y = (int)(qg - (long)x * a.Ny);
                C2 *dst = psi0 + qg * a.Nz + z;
                C2 acc = *dst;
                const T ur = hb ? ar1[q] : ar0[q];
                const T ui = hb ? ai1[q] : ai0[q];
                acc.x += ui;
                acc.y -= ur;
                if (zin_v && x >= a.vb1[0] && x <= a.vb2[0] && y >= a.vb1[1] && y <= a.vb2[1]) {
                    const C2 pn = Ls[line * LS + i1 + m];
                    const long vi = ((long)(x - a.vb1[0]) * (a.vb2[1] - a.vb1[1] + 1) + (y - a.vb1[1])) *
                                        (a.vb2[2] - a.vb1[2] + 1) + (z - a.vb1[2]);
                    const T v = a.vfac * a.vbox[vi];
                    acc.x += v * pn.y;
                    acc.y -= v * pn.x;
                }
                if (zin_box && x >= a.k1a[0] && x <= a.k2b[0] && y >= a.k1a[1] && y <= a.k2b[1] &&
                    !(zin_core && x > a.k1b[0] && x < a.k2a[0] && y > a.k1b[1] && y < a.k2a[1])) {
                    // S = E_x E_y E_z (A_x P_y P_z + P_x A_y P_z + P_x P_y A_z), FP64
                    const double px = a.tfP[0][x], py = a.tfP[1][y], pz = a.tfP[2][z];
                    const double2 ax = a.tfA[0][x], ay = a.tfA[1][y], az = a.tfA[2][z];
                    const double2 ex = a.tfE[0][x], ey = a.tfE[1][y], ez = a.tfE[2][z];
                    const double pyz = py * pz, pxz = px * pz, pxy = px * py;
                    const double cr = ax.x * pyz + ay.x * pxz + az.x * pxy;
                    const double ci = ax.y * pyz + ay.y * pxz + az.y * pxy;
                    const double exy_r = ex.x * ey.x - ex.y * ey.y, exy_i = ex.x * ey.y + ex.y * ey.x;
                    const double e_r = exy_r * ez.x - exy_i * ez.y, e_i = exy_r * ez.y + exy_i * ez.x;
                    const double sr = cr * e_r - ci * e_i, si = cr * e_i + ci * e_r;
                    acc.x -= (T)(sr * a.sin_t - si * a.cos_t);
                    acc.y -= (T)(sr * a.cos_t + si * a.sin_t);
                }
                const T gg = a.g[0][x] * a.g[1][y] * gz;
                acc.x *= gg;
                acc.y *= gg;
                *dst = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Surface capture (AccumulateSurfaceTerms, PSTD_QScat.cpp:1738-2052): per
// surface point, the owning segment's windowed line dotted with the PS and
// FD8 weight vectors, psi at the surface, all times e^{i tau}, accumulated
// in FP64.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_surface(const T *psi_, SurfArgs a) {
    typedef typename VT<T>::c2 C2;
    const SurfFace &F = a.f[blockIdx.y];
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long npts = (long)F.nb * F.nc;
    if (p >= npts) return;
    const long u = p / F.nc, v = p - u * F.nc;
    const C2 *base = reinterpret_cast<const C2 *>(psi_) + u * F.stride_b + v * F.stride_c + F.fft1 * F.stride_l;
    double pr = 0.0, pi = 0.0, fr = 0.0, fi = 0.0;
    for (int l = 0; l < a.L; ++l) {
        const C2 c = base[l * F.stride_l];
        const double w = F.ps[l], w2 = F.fd[l];
        pr += w * (double)c.x;
        pi += w * (double)c.y;
        fr += w2 * (double)c.x;
        fi += w2 * (double)c.y;
    }
    const C2 c = base[(long)F.li * F.stride_l];
    const double cr = (double)c.x, ci = (double)c.y;
    double2 *acc = F.acc;
    double2 s;
    s = acc[p];
    s.x += cr * a.ph_r - ci * a.ph_i;
    s.y += cr * a.ph_i + ci * a.ph_r;
    acc[p] = s;
    s = acc[npts + p];
    s.x += pr * a.ph_r - pi * a.ph_i;
    s.y += pr * a.ph_i + pi * a.ph_r;
    acc[npts + p] = s;
    s = acc[2 * npts + p];
    s.x += fr * a.ph_r - fi * a.ph_i;
    s.y += fr * a.ph_i + fi * a.ph_r;
    acc[2 * npts + p] = s;
}

// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_plane(const PlaneArgs<T> &a, int grid_x, hipStream_t s) {
    dim3 grid(grid_x, a.seg.n), block(64 * a.MB);
    hipLaunchKernelGGL(k_lap_plane<T>, grid, block, plane_lds_bytes<T>(a.MB, a.KB), s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_line(const LineArgs<T> &a, int grid_x, hipStream_t s) {
    dim3 grid(grid_x, a.seg.n), block(64 * a.MB);
    hipLaunchKernelGGL(k_lap_line<T>, grid, block, line_lds_bytes<T>(a.MB, a.KB), s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_surface(const T *psi, const SurfArgs &a, long max_pts, hipStream_t s) {
    dim3 grid((unsigned)((max_pts + 255) / 256), 2), block(256);
    hipLaunchKernelGGL(k_surface<T>, grid, block, 0, s, psi, a);
    return hipGetLastError();
}

hipError_t configure_kernels(int max_lds) {
    hipError_t e;
    if ((e = hipFuncSetAttribute((const void *)k_lap_plane<double>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds))) return e;
    if ((e = hipFuncSetAttribute((const void *)k_lap_plane<float>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds))) return e;
    if ((e = hipFuncSetAttribute((const void *)k_lap_line<double>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds))) return e;
    if ((e = hipFuncSetAttribute((const void *)k_lap_line<float>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds))) return e;
    return hipSuccess;
}

template hipError_t launch_plane<double>(const PlaneArgs<double> &, int, hipStream_t);
template hipError_t launch_plane<float>(const PlaneArgs<float> &, int, hipStream_t);
template hipError_t launch_line<double>(const LineArgs<double> &, int, hipStream_t);
template hipError_t launch_line<float>(const LineArgs<float> &, int, hipStream_t);
template hipError_t launch_surface<double>(const double *, const SurfArgs &, long, hipStream_t);
template hipError_t launch_surface<float>(const float *, const SurfArgs &, long, hipStream_t);
template size_t plane_lds_bytes<double>(int, int);
template size_t plane_lds_bytes<float>(int, int);
template size_t line_lds_bytes<double>(int, int);
template size_t line_lds_bytes<float>(int, int);

}  // namespace pstd
