// kh_kernels.hpp -- gfx950 device kernels of libkhronos_hip.
//
// Field storage: every component lives on one padded grid of
// (nx+2) x (ny+2) x (nz+2) points indexed by the reference's grid index g
// (0 and N+1 are the always-zero boundary planes), x fastest, with the x pitch
// rounded up so that grid index x = 1 of every row is 128-byte aligned.  All
// twelve-plus arrays share one index, so a cell's E, H, D, B, U, W and
// material values are at the same offset.
//
// Update kernel (one per half step, all three components):
//   * a 64 x 4 tile of (x, y) columns per 256-thread block; each thread marches
//     ZB cells along z, carrying the z-neighbour of the differentiated field in
//     registers, so every array element streams from HBM once per half step;
//   * x/y neighbours come from the same or an adjacent tile and hit L1/L2;
//   * tiles are assigned so that consecutive tiles (adjacent in y) land on the
//     same XCD and share its L2;
//   * per cell and component the class is static: cells with a non-zero PML
//     sigma on any axis, or inside a source box, run the reference's full
//     cascade (src/Kernels/Helpers.jl:30-139, :323-340) keeping D/B, U, W;
//     other cells advance E/H directly (FAST) -- or, in EXACT mode, every cell
//     runs the cascade in the reference CPU path's operation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kh {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int MAX_SRC = 8;
constexpr int NXCD = 8;

template <typename T>
struct SrcL {
    int comp;       // 0..2 within the half step
    int s[3];       // first grid index of the box
    int d[3];       // box extent
    const T *amp;   // complex interleaved, column-major over d
    T fr, fi;       // waveform value for this half step
};

template <typename T>
struct HalfArgs {
    T *A[3];              // field being updated (H or E)
    const T *F[3];        // differentiated field (E or H)
    T *Tq[3];             // B or D
    T *U[3];
    T *W[3];
    const T *sig[3];      // sampled PML sigma per axis, indexed by grid index (N+2)
    const T *minv[3];     // per-voxel inverse material, padded layout, or null
    T minv_s;
    T dt;
    T inv[3];
    int n[3];
    int64_t sy, sz;       // element strides of y and z
    int zb;               // z cells per tile
    int tiles[3];         // tile counts in x, y, z
    int per_xcd;          // logical tiles per XCD slot
    int nsrc;
    SrcL<T> src[MAX_SRC];
};

template <typename T>
__device__ __forceinline__ T source_value(const HalfArgs<T> &p, int comp, int gx, int gy, int gz,
                                          bool &inside) {
    // S = sum over sources of Re(f * A) in registration order
    // (src/Sources/Sources.jl:346-357, accumulated into a zeroed S array).
    T s = T(0);
    inside = false;
    for (int q = 0; q < p.nsrc; ++q) {
        const SrcL<T> &o = p.src[q];
        if (o.comp != comp) continue;
        int ix = gx - o.s[0], iy = gy - o.s[1], iz = gz - o.s[2];
        if (ix < 0 || iy < 0 || iz < 0 || ix >= o.d[0] || iy >= o.d[1] || iz >= o.d[2]) continue;
        size_t ai = (size_t)ix + (size_t)o.d[0] * ((size_t)iy + (size_t)o.d[1] * (size_t)iz);
        T ar = o.amp[2 * ai], aim = o.amp[2 * ai + 1];
        T v = (o.fr * ar) - (o.fi * aim);
        s = s + v;
        inside = true;
    }
    return s;
}

// Cell class inside any source box of component `comp` (static over the run).
template <typename T>
__device__ __forceinline__ bool in_source_box(const HalfArgs<T> &p, int comp, int gx, int gy, int gz) {
    for (int q = 0; q < p.nsrc; ++q) {
        const SrcL<T> &o = p.src[q];
        if (o.comp != comp) continue;
        int ix = gx - o.s[0], iy = gy - o.s[1], iz = gz - o.s[2];
        if (ix >= 0 && iy >= 0 && iz >= 0 && ix < o.d[0] && iy < o.d[1] && iz < o.d[2]) return true;
    }
    return false;
}

// One component of one cell.  K is dt * curl in the reference's order.
template <typename T, bool EXACT>
__device__ __forceinline__ void update_component(const HalfArgs<T> &p, int a, size_t idx, T K,
                                                 T so, T sn, T sp, bool pml_cell, int gx, int gy,
                                                 int gz, bool box_cell) {
    T m = p.minv[a] ? p.minv[a][idx] : p.minv_s;
    if (!EXACT && !pml_cell && !box_cell) {
        // E/H advanced directly; D/B are not needed in this cell
        p.A[a][idx] = p.A[a][idx] + m * K;
        return;
    }
    // curl stage: C absent, U present (src/Kernels/Helpers.jl:113-139)
    T tq = p.Tq[a][idx];
    T tn;
    if (sn == T(0)) {
        tn = ((T(1) - sp) * tq + K) / (T(1) + sp);
    } else {
        T uo = p.U[a][idx];
        T un = ((T(1) - sn) * uo + K) / (T(1) + sn);
        p.U[a][idx] = un;
        tn = (((T(1) - sp) * tq + un) - uo) / (T(1) + sp);
    }
    p.Tq[a][idx] = tn;
    // material stage (src/Kernels/Helpers.jl:323-340)
    T net = tn;
    if (box_cell && p.nsrc > 0) {
        bool inside;
        T s = source_value(p, a, gx, gy, gz, inside);
        net = net + s;
    }
    if (so == T(0)) {
        p.A[a][idx] = m * net;
    } else {
        T wo = p.W[a][idx];
        T wn = m * net;
        p.W[a][idx] = wn;
        p.A[a][idx] = (p.A[a][idx] + (T(1) + so) * wn) - (T(1) - so) * wo;
    }
}

__device__ __forceinline__ bool logical_tile(int per_xcd, const int tiles[3], int &tx, int &ty, int &tz) {
    // Physical block b runs on XCD b % 8; give each XCD a contiguous range of
    // logical tiles so that y-neighbouring tiles share an L2.
    int b = blockIdx.x;
    int L = (b % NXCD) * per_xcd + b / NXCD;
    int ntile = tiles[0] * tiles[1] * tiles[2];
    if (L >= ntile) return false;
    tx = L % tiles[0];
    int r = L / tiles[0];
    ty = r % tiles[1];
    tz = r / tiles[1];
    return true;
}

// H half step: K = dt * (d_n2 E_n1 - d_n1 E_n2) with forward differences
// (src/Kernels/Helpers.jl:286-298, idx_curl = +1).
template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void step_h_kernel(HalfArgs<T> p) {
    int tx, ty, tz;
    if (!logical_tile(p.per_xcd, p.tiles, tx, ty, tz)) return;
    const int gx = tx * TX + threadIdx.x + 1;
    const int gy = ty * TY + threadIdx.y + 1;
    if (gx > p.n[0] || gy > p.n[1]) return;
    const int z0 = tz * p.zb + 1;
    const int z1 = min(z0 + p.zb - 1, p.n[2]);
    const int64_t sy = p.sy, sz = p.sz;
    const T *Ex = p.F[0], *Ey = p.F[1], *Ez = p.F[2];
    const T sgx = p.sig[0][gx], sgy = p.sig[1][gy];
    const bool px = sgx != T(0), py = sgy != T(0);
    size_t idx = (size_t)gx + (size_t)gy * sy + (size_t)z0 * sz;
    T ex = Ex[idx], ey = Ey[idx];
    for (int gz = z0; gz <= z1; ++gz, idx += sz) {
        const T sgz = p.sig[2][gz];
        const bool pml = px || py || (sgz != T(0));
        const T ex_z1 = Ex[idx + sz], ey_z1 = Ey[idx + sz];
        const T ez = Ez[idx];
        const T ez_x1 = Ez[idx + 1], ez_y1 = Ez[idx + sy];
        const T ex_y1 = Ex[idx + sy], ey_x1 = Ey[idx + 1];
        // x: d_z Ey - d_y Ez ; y: d_x Ez - d_z Ex ; z: d_y Ex - d_x Ey
        const T Kx = p.dt * (p.inv[2] * (ey_z1 - ey) - p.inv[1] * (ez_y1 - ez));
        const T Ky = p.dt * (p.inv[0] * (ez_x1 - ez) - p.inv[2] * (ex_z1 - ex));
        const T Kz = p.dt * (p.inv[1] * (ex_y1 - ex) - p.inv[0] * (ey_x1 - ey));
        const bool bx = p.nsrc ? in_source_box(p, 0, gx, gy, gz) : false;
        const bool by = p.nsrc ? in_source_box(p, 1, gx, gy, gz) : false;
        const bool bz = p.nsrc ? in_source_box(p, 2, gx, gy, gz) : false;
        // (own, next, prev) = (x,y,z), (y,z,x), (z,x,y)
        update_component<T, EXACT>(p, 0, idx, Kx, sgx, sgy, sgz, pml, gx, gy, gz, bx);
        update_component<T, EXACT>(p, 1, idx, Ky, sgy, sgz, sgx, pml, gx, gy, gz, by);
        update_component<T, EXACT>(p, 2, idx, Kz, sgz, sgx, sgy, pml, gx, gy, gz, bz);
        ex = ex_z1;
        ey = ey_z1;
    }
}

// E half step: backward differences (idx_curl = -1): d(A) = inv * (A[i-1] - A[i]).
template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void step_e_kernel(HalfArgs<T> p) {
    int tx, ty, tz;
    if (!logical_tile(p.per_xcd, p.tiles, tx, ty, tz)) return;
    const int gx = tx * TX + threadIdx.x + 1;
    const int gy = ty * TY + threadIdx.y + 1;
    if (gx > p.n[0] || gy > p.n[1]) return;
    const int z0 = tz * p.zb + 1;
    const int z1 = min(z0 + p.zb - 1, p.n[2]);
    const int64_t sy = p.sy, sz = p.sz;
    const T *Hx = p.F[0], *Hy = p.F[1], *Hz = p.F[2];
    const T sgx = p.sig[0][gx], sgy = p.sig[1][gy];
    const bool px = sgx != T(0), py = sgy != T(0);
    size_t idx = (size_t)gx + (size_t)gy * sy + (size_t)z0 * sz;
    T hx_zm = Hx[idx - sz], hy_zm = Hy[idx - sz];
    for (int gz = z0; gz <= z1; ++gz, idx += sz) {
        const T sgz = p.sig[2][gz];
        const bool pml = px || py || (sgz != T(0));
        const T hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
        const T hz_xm = Hz[idx - 1], hz_ym = Hz[idx - sy];
        const T hx_ym = Hx[idx - sy], hy_xm = Hy[idx - 1];
        // x: d_z Hy - d_y Hz ; y: d_x Hz - d_z Hx ; z: d_y Hx - d_x Hy
        const T Kx = p.dt * (p.inv[2] * (hy_zm - hy) - p.inv[1] * (hz_ym - hz));
        const T Ky = p.dt * (p.inv[0] * (hz_xm - hz) - p.inv[2] * (hx_zm - hx));
        const T Kz = p.dt * (p.inv[1] * (hx_ym - hx) - p.inv[0] * (hy_xm - hy));
        const bool bx = p.nsrc ? in_source_box(p, 0, gx, gy, gz) : false;
        const bool by = p.nsrc ? in_source_box(p, 1, gx, gy, gz) : false;
        const bool bz = p.nsrc ? in_source_box(p, 2, gx, gy, gz) : false;
        update_component<T, EXACT>(p, 0, idx, Kx, sgx, sgy, sgz, pml, gx, gy, gz, bx);
        update_component<T, EXACT>(p, 1, idx, Ky, sgy, sgz, sgx, pml, gx, gy, gz, by);
        update_component<T, EXACT>(p, 2, idx, Kz, sgz, sgx, sgy, pml, gx, gy, gz, bz);
        hx_zm = hx;
        hy_zm = hy;
    }
}

// DFT accumulation of one monitor (src/Monitors/Monitors.jl:361-379):
// acc[cell, k] += (dt * exp(i * f_k * theta)) * field[cell].  The phase
// weights are formed once per block in float64 and kept in LDS; each
// accumulator update is formed in float64 and stored in ACC.
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void dft_kernel(ACC *__restrict__ acc, const T *__restrict__ field,
                                                  int d0, int d1, int d2, int s0, int s1, int s2,
                                                  int64_t sy, int64_t sz, int nf,
                                                  const double *__restrict__ freqs, double theta,
                                                  double dts) {
    extern __shared__ double wsh[];
    for (int k = threadIdx.x; k < nf; k += blockDim.x) {
        double arg = freqs[k] * theta;
        double sn, cs;
        sincos(arg, &sn, &cs);
        wsh[2 * k] = dts * cs;
        wsh[2 * k + 1] = dts * sn;
    }
    __syncthreads();
    const int64_t ncell = (int64_t)d0 * d1 * d2;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    int ix = (int)(cell % d0);
    int64_t r = cell / d0;
    int iy = (int)(r % d1);
    int iz = (int)(r / d1);
    const double v = (double)field[(size_t)(s0 + ix) + (size_t)(s1 + iy) * sy + (size_t)(s2 + iz) * sz];
    ACC *a = acc + 2 * cell;
    for (int k = 0; k < nf; ++k, a += 2 * ncell) {
        double re = (double)a[0] + wsh[2 * k] * v;
        double im = (double)a[1] + wsh[2 * k + 1] * v;
        a[0] = (ACC)re;
        a[1] = (ACC)im;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    // cross-lane butterfly over the 64-wide wavefront
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) s += part[i];
    }
    return s;  // valid in thread 0
}

struct FluxMon {
    const void *acc;  // complex accumulator (dims..., nf)
    int d[3];
};

// One block per frequency: sum over the plane of
// Re(E1 conj(H2) - E2 conj(H1)) * dA after averaging the (up to) two samples
// across the normal axis (src/Monitors/FluxMonitor.jl:104-153).
template <typename ACC>
__global__ __launch_bounds__(256) void flux_kernel(FluxMon e1, FluxMon e2, FluxMon h1, FluxMon h2,
                                                   int ax, int t0, int t1, int n1, int n2, double dA,
                                                   double *out) {
    const int k = blockIdx.x;
    double s = 0.0;
    const FluxMon *m[4] = {&e1, &e2, &h1, &h2};
    for (int c = threadIdx.x; c < n1 * n2; c += blockDim.x) {
        int i1 = c % n1, i2 = c / n1;
        double re[4], im[4];
        for (int q = 0; q < 4; ++q) {
            const FluxMon &M = *m[q];
            const ACC *A = (const ACC *)M.acc;
            int ii[3];
            ii[ax] = 0;
            ii[t0] = i1;
            ii[t1] = i2;
            int64_t nc = (int64_t)M.d[0] * M.d[1] * M.d[2];
            int64_t lin = ii[0] + (int64_t)M.d[0] * (ii[1] + (int64_t)M.d[1] * ii[2]);
            const ACC *p0 = A + 2 * (lin + nc * k);
            double r = (double)p0[0], i = (double)p0[1];
            if (M.d[ax] >= 2) {
                int64_t step = ax == 0 ? 1 : (ax == 1 ? (int64_t)M.d[0] : (int64_t)M.d[0] * M.d[1]);
                const ACC *p1 = p0 + 2 * step;
                r = (r + (double)p1[0]) / 2;
                i = (i + (double)p1[1]) / 2;
            }
            re[q] = r;
            im[q] = i;
        }
        // Re(e1 * conj(h2) - e2 * conj(h1))
        double v = (re[0] * re[3] + im[0] * im[3]) - (re[1] * re[2] + im[1] * im[2]);
        s += v * dA;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) out[k] = s;
}

// Partial sums of |F|^2 over a complex array of n elements.
template <typename ACC>
__global__ __launch_bounds__(256) void sumsq_kernel(const ACC *a, int64_t n, double *partial) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double r = (double)a[2 * i], im = (double)a[2 * i + 1];
        s += r * r + im * im;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

}  // namespace kh
