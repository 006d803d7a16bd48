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
//
// Two variants of each pass:
//   *_reg  operator fragments of the wave's 16-row block live in registers
//          (L <= 96, core <= 128); the psi tile is double-buffered through
//          registers so the next tile's loads overlap this tile's MFMAs.
//   *_lds  operator fragments in LDS; any L whose tiles fit on chip.
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

constexpr int PT = 32;      // complex points per x/y tile (two 16-point blocks)
constexpr int LT = 32;      // lines per z tile (two 16-line blocks)
constexpr int KBMAX = 24;   // register variant: L <= 96
constexpr int NLD = (4 * KBMAX * PT + 255) / 256;   // tile loads per thread (>= 256 threads)

__host__ __device__ constexpr size_t align16(size_t n) { return (n + 15) & ~size_t(15); }

template <typename T> size_t plane_lds_bytes(int MB, int KB) {
    return align16((size_t)MB * KB * 64 * sizeof(T)) + (size_t)4 * KB * PT * 2 * sizeof(T);
}
struct LineInfo;
constexpr size_t LINEINFO_BYTES = 64;
template <typename T> size_t line_lds_bytes(int MB, int KB) {
    return align16((size_t)MB * KB * 64 * sizeof(T)) + align16((size_t)LT * (4 * KB + 1) * 2 * sizeof(T)) +
           LT * LINEINFO_BYTES;
}
template <typename T> size_t plane_reg_lds_bytes(int KB) { return (size_t)4 * KB * PT * 2 * sizeof(T); }
template <typename T> size_t line_reg_lds_bytes(int KB) {
    return align16((size_t)LT * (4 * KB + 1) * 2 * sizeof(T)) + LT * LINEINFO_BYTES;
}
bool reg_variant_ok(int MB, int KB) { return KB <= KBMAX && MB <= 8; }

template <typename C2> __device__ __forceinline__ C2 czero() {
    C2 v;
    v.x = 0;
    v.y = 0;
    return v;
}

// psi0[core row m][col] += -i U for the two 16-point blocks of a x/y tile
template <typename T>
__device__ __forceinline__ void plane_epilogue(const PlaneArgs<T> &a, typename VT<T>::c2 *psi0, int o, long c0,
                                               long core1, int ncore, int wave, int kq, int cq,
                                               const typename VT<T>::acc &ar0, const typename VT<T>::acc &ai0,
                                               const typename VT<T>::acc &ar1, const typename VT<T>::acc &ai1) {
    typedef typename VT<T>::c2 C2;
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

// Per-line terms of the z-pass epilogue, built once per tile in LDS.  With
// the separable TF/SF tables the source is
//   S = E_x E_y E_z ((A_x P_y + P_x A_y) P_z + P_x P_y A_z),
// and the time factor (sin t + i cos t) of PSTD_QScat.cpp:1238-1239 is folded
// into exy.  All FP64.
struct LineInfo {
    double2 exy;    // E_x E_y (sin t + i cos t)
    double2 cxy;    // A_x P_y + P_x A_y
    double pxy;     // P_x P_y
    double gxy;     // g1(x) g2(y)
    long base;      // (x Ny + y) Nz, or -1 past the last line
    int vrow;       // potential-box offset of (x, y, vb1z), or -1
    int flags;      // bit 0: (x, y) in the TF box, bit 1: strictly inside its core
};
static_assert(sizeof(LineInfo) <= LINEINFO_BYTES, "LineInfo must fit its LDS slot");

template <typename T>
__device__ __forceinline__ void line_info(const LineArgs<T> &a, long qg, LineInfo &li) {
    if (qg >= a.nlines) {
        li.base = -1;
        return;
    }
    const int x = (int)qg / a.Ny, y = (int)qg - x * a.Ny;
    li.base = qg * a.Nz;
    li.gxy = (double)a.g[0][x] * (double)a.g[1][y];
    li.vrow = -1;
    if (a.use_v && x >= a.vb1[0] && x <= a.vb2[0] && y >= a.vb1[1] && y <= a.vb2[1])
        li.vrow = ((x - a.vb1[0]) * (a.vb2[1] - a.vb1[1] + 1) + (y - a.vb1[1])) * (a.vb2[2] - a.vb1[2] + 1);
    const bool inbox = x >= a.k1a[0] && x <= a.k2b[0] && y >= a.k1a[1] && y <= a.k2b[1];
    const bool inner = x > a.k1b[0] && x < a.k2a[0] && y > a.k1b[1] && y < a.k2a[1];
    li.flags = (inbox ? 1 : 0) | (inner ? 2 : 0);
    if (inbox) {
        const double px = a.tfP[0][x], py = a.tfP[1][y];
        const double2 ax = a.tfA[0][x], ay = a.tfA[1][y];
        const double2 ex = a.tfE[0][x], ey = a.tfE[1][y];
        li.pxy = px * py;
        li.cxy.x = ax.x * py + ay.x * px;
        li.cxy.y = ax.y * py + ay.y * px;
        const double er = ex.x * ey.x - ex.y * ey.y, ei = ex.x * ey.y + ex.y * ey.x;
        li.exy.x = er * a.sin_t - ei * a.cos_t;
        li.exy.y = er * a.cos_t + ei * a.sin_t;
    }
}

// Per-lane z terms (z is fixed per lane for a block's whole life).
struct ZInfo {
    double2 az, ez;
    double pz, gz;
    int zoff;       // z - vb1z
    int flags;      // bit 0: in potential box, bit 1: in TF box, bit 2: strictly inside
};

template <typename T>
__device__ __forceinline__ ZInfo z_info(const LineArgs<T> &a, int z) {
    ZInfo zi;
    zi.gz = (double)a.g[2][z];
    zi.zoff = z - a.vb1[2];
    const bool inv = a.use_v && z >= a.vb1[2] && z <= a.vb2[2];
    const bool inbox = z >= a.k1a[2] && z <= a.k2b[2];
    const bool inner = z > a.k1b[2] && z < a.k2a[2];
    zi.flags = (inv ? 1 : 0) | (inbox ? 2 : 0) | (inner ? 4 : 0);
    zi.pz = a.tfP[2][z];
    zi.az = a.tfA[2][z];
    zi.ez = a.tfE[2][z];
    return zi;
}

// z-pass epilogue: psi0 = g (psi0 - i U - i f V~ psi - S i e^{-it}), the
// potential with its ramp (PSTD_QScat.cpp:1268-1303), TF/SF source
// (:1306-1419, regenerated from separable FP64 tables and the FP64 time
// phase) and the ABC mask (:1421-1435).
template <typename T>
__device__ __forceinline__ void line_epilogue(const LineArgs<T> &a, const typename VT<T>::c2 *Ls, const LineInfo *LI,
                                              int LS, const ZInfo &zi, int z, int m, int i1, int kq,
                                              const typename VT<T>::acc &ar0, const typename VT<T>::acc &ai0,
                                              const typename VT<T>::acc &ar1, const typename VT<T>::acc &ai1) {
    typedef typename VT<T>::c2 C2;
    C2 *psi0 = reinterpret_cast<C2 *>(a.psi0);
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int line = hb * 16 + VT<T>::row(kq, q);
            const LineInfo &li = LI[line];
            if (li.base < 0) continue;
            C2 *dst = psi0 + li.base + z;
            C2 acc = *dst;
            const T ur = hb ? ar1[q] : ar0[q];
            const T ui = hb ? ai1[q] : ai0[q];
            acc.x += ui;
            acc.y -= ur;
            if ((zi.flags & 1) && li.vrow >= 0) {
                const C2 pn = Ls[line * LS + i1 + m];
                const T v = a.vfac * a.vbox[li.vrow + zi.zoff];
                acc.x += v * pn.y;
                acc.y -= v * pn.x;
            }
            if ((zi.flags & 2) && (li.flags & 1) && !((zi.flags & 4) && (li.flags & 2))) {
                const double ur2 = li.cxy.x * zi.pz + li.pxy * zi.az.x;
                const double ui2 = li.cxy.y * zi.pz + li.pxy * zi.az.y;
                const double er = li.exy.x * zi.ez.x - li.exy.y * zi.ez.y;
                const double ei = li.exy.x * zi.ez.y + li.exy.y * zi.ez.x;
                acc.x -= (T)(er * ur2 - ei * ui2);
                acc.y -= (T)(er * ui2 + ei * ur2);
            }
            const T gg = (T)(li.gxy * zi.gz);
            acc.x *= gg;
            acc.y *= gg;
            *dst = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// x / y pass, register variant.  Block = one segment (blockIdx.y) and a
// strided set of column tiles; wave w owns 16-row block w of the core and
// keeps its operator fragments in registers for the life of the block.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(512) k_lap_plane_reg(PlaneArgs<T> a) {
    typedef typename VT<T>::c2 C2;
    typedef typename VT<T>::acc ACC;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    C2 *Bs = reinterpret_cast<C2 *>(smem);
    const int r = blockIdx.y;
    const int KB = a.KB, Kp = 4 * KB;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mbr = a.seg.mb[r];
    const bool active = wave < mbr;
    const T *op = a.op + a.seg.opoff[r] + (long)(active ? wave : 0) * KB * 64 + lane;
    T af[KBMAX];
#pragma unroll
    for (int kb = 0; kb < KBMAX; ++kb) af[kb] = (active && kb < KB) ? op[kb * 64] : T(0);
    const long fft1 = a.seg.fft1[r], core1 = a.seg.core1[r];
    const int ncore = a.seg.ncore[r];
    const int ntiles = a.outer_count * a.col_tiles;
    const int nel = Kp * PT;
    const C2 *psi = reinterpret_cast<const C2 *>(a.psi);
    C2 *psi0 = reinterpret_cast<C2 *>(a.psi0);
    const int kq = lane >> 4, cq = lane & 15;
    C2 pre[NLD];
    auto fetch = [&](int t) {
        const int o = t / a.col_tiles;
        const long c0 = (long)(t - o * a.col_tiles) * PT;
        const C2 *src = psi + (long)o * a.outer_stride + fft1 * a.row_stride + c0;
#pragma unroll
        for (int j = 0; j < NLD; ++j) {
            const int i = tid + j * nthr;
            const int l = i / PT, c = i % PT;
            C2 v = czero<C2>();
            if (i < nel && l < a.L && c0 + c < a.ncols) v = src[(long)l * a.row_stride + c];
            pre[j] = v;
        }
    };
    int t = blockIdx.x;
    if (t < ntiles) fetch(t);
    for (; t < ntiles; t += gridDim.x) {
        const int o = t / a.col_tiles;
        const long c0 = (long)(t - o * a.col_tiles) * PT;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NLD; ++j) {
            const int i = tid + j * nthr;
            if (i < nel) Bs[i] = pre[j];
        }
        __syncthreads();
        if (t + (int)gridDim.x < ntiles) fetch(t + gridDim.x);
        if (active) {
            ACC ar0 = {0, 0, 0, 0}, ai0 = {0, 0, 0, 0}, ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
            const C2 *Bw = Bs + kq * PT + cq;
#pragma unroll
            for (int kb = 0; kb < KBMAX; ++kb) {
                if (kb < KB) {
                    const C2 b0 = Bw[kb * 4 * PT];
                    const C2 b1 = Bw[kb * 4 * PT + 16];
                    ar0 = mfma16(af[kb], b0.x, ar0);
                    ai0 = mfma16(af[kb], b0.y, ai0);
                    ar1 = mfma16(af[kb], b1.x, ar1);
                    ai1 = mfma16(af[kb], b1.y, ai1);
                }
            }
            plane_epilogue<T>(a, psi0, o, c0, core1, ncore, wave, kq, cq, ar0, ai0, ar1, ai1);
        }
    }
}

// ---------------------------------------------------------------------------
// z pass (contiguous lines), register variant, with the pointwise terms of
// Update fused into the epilogue.  The MFMA roles are swapped (psi is the A
// operand, the operator the B operand) so that each lane ends with both
// components of one output point and 16 lanes cover 16 consecutive z.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(512) k_lap_line_reg(LineArgs<T> a) {
    typedef typename VT<T>::c2 C2;
    typedef typename VT<T>::acc ACC;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int r = blockIdx.y;
    const int KB = a.KB, Kp = 4 * KB, LS = Kp + 1;
    C2 *Ls = reinterpret_cast<C2 *>(smem);
    LineInfo *LI = reinterpret_cast<LineInfo *>(smem + align16((size_t)LT * LS * sizeof(C2)));
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mbr = a.seg.mb[r];
    const long fft1 = a.seg.fft1[r];
    const int core1 = a.seg.core1[r], ncore = a.seg.ncore[r];
    const int i1 = core1 - (int)fft1;
    const int kq = lane >> 4, cq = lane & 15;
    const int m = wave * 16 + cq;
    // a wave is active if its 16-row block exists; a lane writes if its row
    // is a core row
    const bool active = wave < mbr;
    const bool lane_ok = active && m < ncore;
    const int z = core1 + (lane_ok ? m : 0);
    const ZInfo zi = z_info<T>(a, z);
    const T *op = a.op + a.seg.opoff[r] + (long)(active ? wave : 0) * KB * 64 + lane;
    T af[KBMAX];
#pragma unroll
    for (int kb = 0; kb < KBMAX; ++kb) af[kb] = (active && kb < KB) ? op[kb * 64] : T(0);
    const int nel = LT * Kp;
    const C2 *psi = reinterpret_cast<const C2 *>(a.psi);
    C2 pre[NLD];
    auto fetch = [&](int t) {
        const long q0 = (long)t * LT;
#pragma unroll
        for (int j = 0; j < NLD; ++j) {
            const int i = tid + j * nthr;
            const int line = i / Kp, l = i - line * Kp;
            C2 v = czero<C2>();
            if (i < nel && l < a.L && q0 + line < a.nlines) v = psi[(q0 + line) * a.Nz + fft1 + l];
            pre[j] = v;
        }
    };
    int t = blockIdx.x;
    if (t < a.line_tiles) fetch(t);
    for (; t < a.line_tiles; t += gridDim.x) {
        const long q0 = (long)t * LT;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NLD; ++j) {
            const int i = tid + j * nthr;
            if (i < nel) {
                const int line = i / Kp, l = i - line * Kp;
                Ls[line * LS + l] = pre[j];
            }
        }
        if (tid < LT) line_info<T>(a, q0 + tid, LI[tid]);
        __syncthreads();
        if (t + (int)gridDim.x < a.line_tiles) fetch(t + gridDim.x);
        if (!active) continue;
        ACC ar0 = {0, 0, 0, 0}, ai0 = {0, 0, 0, 0}, ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
        const C2 *L0 = Ls + cq * LS + kq;
        const C2 *L1 = Ls + (16 + cq) * LS + kq;
#pragma unroll
        for (int kb = 0; kb < KBMAX; ++kb) {
            if (kb < KB) {
                const C2 p0 = L0[4 * kb];
                const C2 p1 = L1[4 * kb];
                ar0 = mfma16(p0.x, af[kb], ar0);
                ai0 = mfma16(p0.y, af[kb], ai0);
                ar1 = mfma16(p1.x, af[kb], ar1);
                ai1 = mfma16(p1.y, af[kb], ai1);
            }
        }
        if (lane_ok) line_epilogue<T>(a, Ls, LI, LS, zi, z, m, i1, kq, ar0, ai0, ar1, ai1);
    }
}

// ---------------------------------------------------------------------------
// LDS-operator variants (any L whose operator and tile fit in LDS).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_lap_plane_lds(PlaneArgs<T> a) {
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
            C2 v = czero<C2>();
            if (l < a.L && c0 + c < a.ncols) v = src[(long)l * a.row_stride + c];
            Bs[i] = v;
        }
        __syncthreads();
        if (wave < mbr) {
            ACC ar0 = {0, 0, 0, 0}, ai0 = {0, 0, 0, 0}, ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
            const T *Aw = As + wave * KB * 64 + lane;
            const C2 *Bw = Bs + kq * PT + cq;
            for (int kb = 0; kb < KB; ++kb) {
                const T af = Aw[kb * 64];
                const C2 b0 = Bw[kb * 4 * PT];
                const C2 b1 = Bw[kb * 4 * PT + 16];
                ar0 = mfma16(af, b0.x, ar0);
                ai0 = mfma16(af, b0.y, ai0);
                ar1 = mfma16(af, b1.x, ar1);
                ai1 = mfma16(af, b1.y, ai1);
            }
            plane_epilogue<T>(a, psi0, o, c0, core1, ncore, wave, kq, cq, ar0, ai0, ar1, ai1);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_lap_line_lds(LineArgs<T> a) {
    typedef typename VT<T>::c2 C2;
    typedef typename VT<T>::acc ACC;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int r = blockIdx.y;
    const int KB = a.KB, Kp = 4 * KB, LS = Kp + 1;
    T *As = reinterpret_cast<T *>(smem);
    const size_t a_bytes = align16((size_t)a.MB * KB * 64 * sizeof(T));
    C2 *Ls = reinterpret_cast<C2 *>(smem + a_bytes);
    LineInfo *LI = reinterpret_cast<LineInfo *>(smem + a_bytes + align16((size_t)LT * LS * sizeof(C2)));
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mbr = a.seg.mb[r];
    const T *op = a.op + a.seg.opoff[r];
    for (int i = tid; i < mbr * KB * 64; i += nthr) As[i] = op[i];
    const long fft1 = a.seg.fft1[r];
    const int core1 = a.seg.core1[r], ncore = a.seg.ncore[r];
    const int i1 = core1 - (int)fft1;
    const C2 *psi = reinterpret_cast<const C2 *>(a.psi);
    const int kq = lane >> 4, cq = lane & 15;
    const int m = wave * 16 + cq;
    const bool lane_ok = wave < mbr && m < ncore;
    const int z = core1 + (lane_ok ? m : 0);
    const ZInfo zi = z_info<T>(a, z);
    for (int t = blockIdx.x; t < a.line_tiles; t += gridDim.x) {
        const long q0 = (long)t * LT;
        __syncthreads();
        for (int i = tid; i < LT * Kp; i += nthr) {
            const int line = i / Kp, l = i - line * Kp;
            C2 v = czero<C2>();
            if (l < a.L && q0 + line < a.nlines) v = psi[(q0 + line) * a.Nz + fft1 + l];
            Ls[line * LS + l] = v;
        }
        if (tid < LT) line_info<T>(a, q0 + tid, LI[tid]);
        __syncthreads();
        if (wave >= mbr) continue;
        ACC ar0 = {0, 0, 0, 0}, ai0 = {0, 0, 0, 0}, ar1 = {0, 0, 0, 0}, ai1 = {0, 0, 0, 0};
        const T *Aw = As + wave * KB * 64 + lane;
        const C2 *L0 = Ls + cq * LS + kq;
        const C2 *L1 = Ls + (16 + cq) * LS + kq;
        for (int kb = 0; kb < KB; ++kb) {
            const T af = Aw[kb * 64];
            const C2 p0 = L0[4 * kb];
            const C2 p1 = L1[4 * kb];
            ar0 = mfma16(p0.x, af, ar0);
            ai0 = mfma16(p0.y, af, ai0);
            ar1 = mfma16(p1.x, af, ar1);
            ai1 = mfma16(p1.y, af, ai1);
        }
        if (lane_ok) line_epilogue<T>(a, Ls, LI, LS, zi, z, m, i1, kq, ar0, ai0, ar1, ai1);
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
    dim3 grid(grid_x, a.seg.n);
    if (reg_variant_ok(a.MB, a.KB)) {
        dim3 block(64 * (a.MB > 4 ? a.MB : 4));
        hipLaunchKernelGGL(k_lap_plane_reg<T>, grid, block, plane_reg_lds_bytes<T>(a.KB), s, a);
    } else {
        dim3 block(64 * a.MB);
        hipLaunchKernelGGL(k_lap_plane_lds<T>, grid, block, plane_lds_bytes<T>(a.MB, a.KB), s, a);
    }
    return hipGetLastError();
}
template <typename T>
hipError_t launch_line(const LineArgs<T> &a, int grid_x, hipStream_t s) {
    dim3 grid(grid_x, a.seg.n);
    if (reg_variant_ok(a.MB, a.KB)) {
        dim3 block(64 * (a.MB > 4 ? a.MB : 4));
        hipLaunchKernelGGL(k_lap_line_reg<T>, grid, block, line_reg_lds_bytes<T>(a.KB), s, a);
    } else {
        dim3 block(64 * a.MB);
        hipLaunchKernelGGL(k_lap_line_lds<T>, grid, block, line_lds_bytes<T>(a.MB, a.KB), s, a);
    }
    return hipGetLastError();
}
template <typename T>
hipError_t launch_surface(const T *psi, const SurfArgs &a, long max_pts, hipStream_t s) {
    dim3 grid((unsigned)((max_pts + 255) / 256), 2), block(256);
    hipLaunchKernelGGL(k_surface<T>, grid, block, 0, s, psi, a);
    return hipGetLastError();
}

// Resident blocks per CU of the pass kernel that launch_plane / launch_line
// would pick for (MB, KB), from the runtime's occupancy calculator.
template <typename T> int pass_occupancy(int MB, int KB, bool line) {
    int n = 0;
    hipError_t e;
    if (reg_variant_ok(MB, KB)) {
        const int block = 64 * (MB > 4 ? MB : 4);
        if (line)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_lap_line_reg<T>, block, line_reg_lds_bytes<T>(KB));
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_lap_plane_reg<T>, block, plane_reg_lds_bytes<T>(KB));
    } else {
        const int block = 64 * MB;
        if (line)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_lap_line_lds<T>, block, line_lds_bytes<T>(MB, KB));
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_lap_plane_lds<T>, block, plane_lds_bytes<T>(MB, KB));
    }
    return (e == hipSuccess && n > 0) ? n : 1;
}

template <typename T> size_t pass_lds_bytes(int MB, int KB) {
    if (reg_variant_ok(MB, KB)) return std::max(plane_reg_lds_bytes<T>(KB), line_reg_lds_bytes<T>(KB));
    return std::max(plane_lds_bytes<T>(MB, KB), line_lds_bytes<T>(MB, KB));
}

hipError_t configure_kernels(int max_lds) {
    const void *k[] = {(const void *)k_lap_plane_reg<double>, (const void *)k_lap_plane_reg<float>,
                       (const void *)k_lap_line_reg<double>,  (const void *)k_lap_line_reg<float>,
                       (const void *)k_lap_plane_lds<double>, (const void *)k_lap_plane_lds<float>,
                       (const void *)k_lap_line_lds<double>,  (const void *)k_lap_line_lds<float>};
    for (const void *f : k) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template hipError_t launch_plane<double>(const PlaneArgs<double> &, int, hipStream_t);
template hipError_t launch_plane<float>(const PlaneArgs<float> &, int, hipStream_t);
template hipError_t launch_line<double>(const LineArgs<double> &, int, hipStream_t);
template hipError_t launch_line<float>(const LineArgs<float> &, int, hipStream_t);
template hipError_t launch_surface<double>(const double *, const SurfArgs &, long, hipStream_t);
template hipError_t launch_surface<float>(const float *, const SurfArgs &, long, hipStream_t);
template size_t pass_lds_bytes<double>(int, int);
template size_t pass_lds_bytes<float>(int, int);

}  // namespace pstd

namespace pstd {
template int pass_occupancy<double>(int, int, bool);
template int pass_occupancy<float>(int, int, bool);
}  // namespace pstd
