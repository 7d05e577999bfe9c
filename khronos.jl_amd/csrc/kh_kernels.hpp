// kh_kernels.hpp -- gfx950 device kernels of libkhronos_hip.
//
// Field storage: every component lives on one padded grid of
// (nx+2) x (ny+2) x (nz+2) points indexed by the reference's grid index g
// (0 and N+1 are the always-zero boundary planes), x fastest, with the x pitch
// rounded up so that grid index x = 1 of every row is 128-byte aligned.  All
// arrays share one index, so a cell's E, H, D, B, U, W and material values sit
// at the same offset.
//
// A half step (H or E) is split by cell class:
//
//   inner_*_kernel  -- FAST mode, the box of cells whose PML sigma is zero on
//     all three axes.  E/H advance directly (E += eps^-1 * dt * curl H, the
//     D/B-eliminated form of src/Kernels/CUDAKernels.jl:46-74).  Each lane owns
//     two adjacent x cells (one 8/16-byte access per array) of ZB planes; with
//     ZB > 1 the loads of plane z+1 are issued before the stores of plane z.
//     Streams touched once (updated field, material) use nontemporal access.
//
//   pml_*_kernel -- FAST mode, the six PML slabs around the inner box: the
//     sigma cascade with D/B eliminated (carried on E/H or W), x pairs per lane.
//
//   cascade_*_kernel -- EXACT mode, the whole grid, one cell per thread: the
//     reference's sigma cascade keeping D/B, U and W (src/Kernels/Helpers.jl:
//     113-139, :323-340) in the CPU path's operation order, with current
//     sources and the total polarization in the material stage.
//
// After the E update: source_fix_kernel (FAST: the terms the eliminated D/B
// would have carried), wrap_kernel (periodic / Bloch ghosts), material_kernel
// (chi3 and dispersive poles); then dft_kernel for the due monitors of each
// half step and tick_kernel, which advances the device timestep every
// per-step quantity is computed from.  Complex fields run the real kernels on
// the real and imaginary parts in turn.
//
// Blocks are assigned so that consecutive logical blocks (neighbouring rows
// and planes) run on the same XCD and share its L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kh {

constexpr int TX = 64;       // lanes per inner-tile row
constexpr int TY = 4;        // rows per inner tile
constexpr int MAX_SRC = 8;
constexpr int MAX_DFT = 16;   // DFT monitors per half step (one launch)
constexpr int MAX_DFTE = 4;   // small E monitors accumulated by the E update itself
constexpr int MAX_SLAB = 6;
constexpr int NXCD = 8;
constexpr int CBLK = 256;    // threads per cascade block

template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };

template <typename T> struct EV2;
template <> struct EV2<float> { typedef float t __attribute__((ext_vector_type(2))); };
template <> struct EV2<double> { typedef double t __attribute__((ext_vector_type(2))); };

// The timestep lives in device memory and is advanced by tick_kernel at the
// end of every step, so no launch of a step carries per-step arguments: the
// time of a half step, the waveform values and the DFT gates are read from it
// (a block of steps can then be captured once as a graph and replayed).
template <typename T>
__device__ __forceinline__ double step_time(int64_t ts, T dt) {
    return (double)((T)ts * dt);  // src/Simulation.jl:22
}

// Sources are injected while t <= cutoff (src/Kernels/Kernels.jl:27-35); t
// is non-decreasing in ts, so the host's sticky switch-off is this test.
template <typename T>
__device__ __forceinline__ bool sources_on(int64_t ts, T dt, double cutoff) {
    return ts >= 0 && !(step_time(ts, dt) > cutoff);
}

// Waveform value of step n from the caller's table (zero past it).
template <typename T>
__device__ __forceinline__ void wave_value(const double *w, int64_t nw, int64_t n, T &fr, T &fi) {
    fr = T(0);
    fi = T(0);
    if (n < 0 || n >= nw) return;
    fr = (T)w[2 * n];
    fi = (T)w[2 * n + 1];
}

template <typename T>
struct SrcL {
    int comp;          // 0..2 within the half step
    int s[3];          // first grid index of the box
    int d[3];          // box extent
    const T *amp;      // complex interleaved, column-major over d
    const double *wave;  // (re, im) per step of this half step
    int64_t nwave;
};

// A current source as the FAST correction applies it (see source_fix_kernel).
template <typename T>
struct FixSrc {
    T *A;
    T *W;
    const T *minv;        // per-voxel or null (minv_s)
    const T *sig_own, *sig_prev;
    int own, prev;
    const T *amp;
    int s[3], d[3];
    const double *wave;
    int64_t nwave;
    int first;            // first block
};

struct Slab {
    int lo[3];      // first grid index
    int n[3];       // extent
    int first;      // first logical block of the slab
};

template <typename T>
struct CascadeArgs {
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
    int64_t sy, sz;       // element strides of y and z
    int nslab;
    Slab slab[MAX_SLAB];
    int nblock;           // logical blocks over all slabs
    int per_xcd;
    int nsrc;
    const SrcL<T> *src;   // device array of nsrc sources of this half step, registration order
    // material conductivity (per voxel, padded grid) and the C stage
    // arrays; null for a component without conductivity
    const T *msig[3];
    T *C[3];
    int noU[3];           // component has no U stage (its next-axis profile is absent)
    const int64_t *clk;   // device timestep of the first step of the launch block
    int koff;             // this step's offset in the block (graph position)
    double cutoff;        // last source time
    // total polarization subtracted in the material stage (E, EXACT mode):
    // dense over the material box ulo..ulo+un-1, or null
    const T *fpd[3];
    int ulo[3], un[3];
};

// One dispersive pole on the device: coefficients in the backend type and
// its arrays, dense over the pole's box (x fastest); [part][component], the
// imaginary part only with complex fields.
template <typename T>
struct PoleDev {
    T g1i, g1;        // gamma1_inv, gamma1
    T coefp;          // 2 - omega0_dt_sq (Lorentz) or 2 (Drude)
    T drive;          // sigma_omega0_dt_sq (Lorentz) or drude_coeff (Drude)
    int lo[3], n[3];  // box: first grid index and extent
    const T *sigma;
    T *P[2][3];
    T *Pp[2][3];
};

// Material pass after the E update: chi3 and the poles, over the material box
// (bounding box of chi3 != 0 and every pole's box), one cell per thread.
// Arrays are [part][component] (part 1: imaginary, complex fields only).
template <typename T>
struct MatArgs {
    T *E[2][3];
    T *W[2][3];           // E-side own-axis stage (FAST correction in the PML)
    const T *minv[3];
    T minv_s;
    const T *sig[3];      // sampled D-side PML sigma per axis (N+2)
    const T *chi3;        // dense over the box, or null
    T *delta[2][3];       // FAST with chi3: linear minus stored E per cell, or null
    // total polarization, two buffers (FAST: alternating by timestep parity,
    // buffer ts % 2 holds the total the update used and buffer 1 - ts % 2 the
    // previous one, replaced by the new total; EXACT: buffer 0 only)
    T *fpd[2][2][3];      // [buffer][part][component]; null without poles
    const int64_t *clk;
    int koff;
    int lo[3], n[3];
    int64_t sy, sz;
    int npole;
    const PoleDev<T> *poles;
    // per box cell: 1 where some pole's sigma or chi3 is non-zero, or null.
    // Elsewhere P stays zero (it is only updated where its sigma is not), so
    // fPD, the FAST corrections and the chi3 delta are zero and the pass would
    // store what it read: such cells are skipped.
    const uint8_t *active;
    // FAST with a periodic axis: the wrapped ghost of a corrected boundary
    // cell is rewritten from the corrected linear field
    int N[3];
    int per[3];
    double fwd[3][2], rev[3][2];
};

template <typename T>
struct InnerArgs {
    T *A[3];
    const T *F[3];
    const T *minv[3];     // per-voxel inverse material or null (scalar minv_s)
    T minv_s;
    T dt;
    T inv[3];
    int lo[3], hi[3];     // the inner box, inclusive grid indices
    int x0;               // grid x of the first tile's first cell (1 + 128k)
    int64_t sy, sz;
    int zb;
    int tiles[3];
    int per_xcd;
};

// Physical block b runs on XCD b % 8.  Each XCD takes pairs of consecutive
// logical blocks in turn (XCD x, its j-th block: logical pair 8 (j / 2) + x),
// so the eight XCDs stream the same stretch of the grid together and a pair
// of neighbouring row segments shares an L2.  Measured on the strip
// waveguide's PML updates: 7% faster than one contiguous range of logical
// blocks per XCD, 1% faster than plain round robin.  per_xcd (physical blocks
// per XCD) comes from xcd_per() on the host, which covers every logical block.
constexpr int XCD_PAIR = 2;

__device__ __forceinline__ bool xcd_block(int per_xcd, int nblock, int &L) {
    const int b = blockIdx.x;
    const int x = b % NXCD, j = b / NXCD;
    L = ((j / XCD_PAIR) * NXCD + x) * XCD_PAIR + j % XCD_PAIR;
    return j < per_xcd && L < nblock;
}

// physical blocks per XCD so that xcd_block covers logical blocks 0..nblock-1
inline int xcd_per(int nblock) {
    const int pairs = (nblock + XCD_PAIR - 1) / XCD_PAIR;
    return (pairs + NXCD - 1) / NXCD * XCD_PAIR;
}

// ------------------------------------------------------- memory helpers
template <typename T>
__device__ __forceinline__ typename V2<T>::t ld2(const T *p, size_t i) {
    return *reinterpret_cast<const typename V2<T>::t *>(p + i);
}

template <typename T>
__device__ __forceinline__ void st2(T *p, size_t i, typename V2<T>::t v) {
    *reinterpret_cast<typename V2<T>::t *>(p + i) = v;
}

// Streams touched once per half step may bypass cache retention (NT =
// nontemporal); the differentiated field is re-read by neighbouring lanes,
// rows and planes and keeps normal caching.
template <typename T, bool NT>
__device__ __forceinline__ typename V2<T>::t ldo(const T *p, size_t i) {
    if constexpr (NT) {
        typename EV2<T>::t v = __builtin_nontemporal_load(reinterpret_cast<const typename EV2<T>::t *>(p + i));
        typename V2<T>::t r;
        r.x = v.x;
        r.y = v.y;
        return r;
    } else {
        return ld2(p, i);
    }
}

template <typename T, bool NT>
__device__ __forceinline__ void sto(T *p, size_t i, typename V2<T>::t v) {
    if constexpr (NT) {
        typename EV2<T>::t w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, reinterpret_cast<typename EV2<T>::t *>(p + i));
    } else {
        st2(p, i, v);
    }
}

template <typename T>
__device__ __forceinline__ T ld1(const T *p, size_t i) {
    return __builtin_nontemporal_load(p + i);
}

template <typename T>
__device__ __forceinline__ void st1(T *p, size_t i, T v) {
    __builtin_nontemporal_store(v, p + i);
}

// ------------------------------------------------------- cascade kernels
template <typename T>
__device__ __forceinline__ T source_value(const CascadeArgs<T> &p, int comp, int gx, int gy, int gz,
                                          bool &inside) {
    // S = sum over sources of Re(f * A) in registration order
    // (src/Sources/Sources.jl:346-357, accumulated into a zeroed S array).
    T s = T(0);
    inside = false;
    const int64_t ts = *p.clk + p.koff;
    const bool on = sources_on(ts, p.dt, p.cutoff);
    for (int q = 0; q < p.nsrc; ++q) {
        const SrcL<T> &o = p.src[q];
        if (o.comp != comp) continue;
        int ix = gx - o.s[0], iy = gy - o.s[1], iz = gz - o.s[2];
        if (ix < 0 || iy < 0 || iz < 0 || ix >= o.d[0] || iy >= o.d[1] || iz >= o.d[2]) continue;
        T fr = T(0), fi = T(0);
        if (on) wave_value(o.wave, o.nwave, ts, fr, fi);
        size_t ai = (size_t)ix + (size_t)o.d[0] * ((size_t)iy + (size_t)o.d[1] * (size_t)iz);
        T ar = o.amp[2 * ai], aim = o.amp[2 * ai + 1];
        T v = (fr * ar) - (fi * aim);
        s = s + v;
        inside = true;
    }
    return s;
}

// One component of one cell, reference cascade.  K is dt * curl in the
// reference's order; so, sn, sp are the own/next/prev-axis sigma samples.
template <typename T>
__device__ __forceinline__ void cascade_component(const CascadeArgs<T> &p, int a, size_t idx, T K, T so,
                                                  T sn, T sp, int gx, int gy, int gz) {
    const T m = p.minv[a] ? ld1(p.minv[a], idx) : p.minv_s;
    const T tq = ld1((const T *)p.Tq[a], idx);
    T tn;
    if (p.C[a] && p.noU[a]) {
        // C allocated, no U (2D, next axis z; src/Kernels/Helpers.jl:80-93)
        const T sd = T(0.5) * (p.dt * ld1(p.msig[a], idx));
        if (sd == T(0)) {
            tn = ((T(1) - sp) * tq + K) / (T(1) + sp);
        } else {
            const T co = ld1((const T *)p.C[a], idx);
            const T cn = ((T(1) - sd) * co + K) / (T(1) + sd);
            st1(p.C[a], idx, cn);
            tn = (((T(1) - sp) * tq + cn) - co) / (T(1) + sp);
        }
    } else if (p.C[a]) {
        // curl stage with C and U allocated and a conductivity array
        // (src/Kernels/Helpers.jl:39-70): sigma~ = 0.5 (dt sigma_D)
        // (:278-279); sigma~ = 0 skips C, and U is never bypassed here
        const T sd = T(0.5) * (p.dt * ld1(p.msig[a], idx));
        const T uo = ld1((const T *)p.U[a], idx);
        T un;
        if (sd == T(0)) {
            un = ((T(1) - sn) * uo + K) / (T(1) + sn);
        } else {
            const T co = ld1((const T *)p.C[a], idx);
            const T cn = ((T(1) - sd) * co + K) / (T(1) + sd);
            st1(p.C[a], idx, cn);
            un = (((T(1) - sn) * uo + cn) - co) / (T(1) + sn);
        }
        st1(p.U[a], idx, un);
        tn = (((T(1) - sp) * tq + un) - uo) / (T(1) + sp);
    } else if (sn == T(0)) {
        // curl stage: C absent, U present (src/Kernels/Helpers.jl:113-139)
        tn = ((T(1) - sp) * tq + K) / (T(1) + sp);
    } else {
        const T uo = ld1((const T *)p.U[a], idx);
        const T un = ((T(1) - sn) * uo + K) / (T(1) + sn);
        st1(p.U[a], idx, un);
        tn = (((T(1) - sp) * tq + un) - uo) / (T(1) + sp);
    }
    st1(p.Tq[a], idx, tn);
    // material stage (src/Kernels/Helpers.jl:323-340)
    T net = tn;
    if (p.nsrc > 0) {
        bool inside;
        const T s = source_value(p, a, gx, gy, gz, inside);
        if (inside) net = net + s;
    }
    if (p.fpd[0]) {
        // net -= fPD (src/Kernels/Helpers.jl:330-333); fPD is zero outside the box
        const int ux = gx - p.ulo[0], uy = gy - p.ulo[1], uz = gz - p.ulo[2];
        if (ux >= 0 && uy >= 0 && uz >= 0 && ux < p.un[0] && uy < p.un[1] && uz < p.un[2])
            net = net - p.fpd[a][(size_t)ux + (size_t)p.un[0] * ((size_t)uy + (size_t)p.un[1] * (size_t)uz)];
    }
    if (so == T(0)) {
        st1(p.A[a], idx, m * net);
    } else {
        const T wo = ld1((const T *)p.W[a], idx);
        const T ao = ld1((const T *)p.A[a], idx);
        const T wn = m * net;
        st1(p.W[a], idx, wn);
        st1(p.A[a], idx, (ao + (T(1) + so) * wn) - (T(1) - so) * wo);
    }
}

template <typename T>
__device__ __forceinline__ bool cascade_cell(const CascadeArgs<T> &p, int &gx, int &gy, int &gz) {
    int L;
    if (!xcd_block(p.per_xcd, p.nblock, L)) return false;
    int s = 0;
    while (s + 1 < p.nslab && L >= p.slab[s + 1].first) ++s;
    const Slab &b = p.slab[s];
    const int64_t c = (int64_t)(L - b.first) * CBLK + threadIdx.x;
    const int64_t ncell = (int64_t)b.n[0] * b.n[1] * b.n[2];
    if (c >= ncell) return false;
    gx = b.lo[0] + (int)(c % b.n[0]);
    const int64_t r = c / b.n[0];
    gy = b.lo[1] + (int)(r % b.n[1]);
    gz = b.lo[2] + (int)(r / b.n[1]);
    return true;
}

// H half step, cascade: K = dt * (d_n2 E_n1 - d_n1 E_n2) with forward
// differences (src/Kernels/Helpers.jl:286-298, idx_curl = +1).
template <typename T>
__global__ __launch_bounds__(CBLK) void cascade_h_kernel(CascadeArgs<T> p) {
    int gx, gy, gz;
    if (!cascade_cell(p, gx, gy, gz)) return;
    const int64_t sy = p.sy, sz = p.sz;
    const size_t idx = (size_t)gx + (size_t)gy * sy + (size_t)gz * sz;
    const T sgx = p.sig[0][gx], sgy = p.sig[1][gy], sgz = p.sig[2][gz];
    const T *Ex = p.F[0], *Ey = p.F[1], *Ez = p.F[2];
    const T ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
    const T ex_z1 = Ex[idx + sz], ey_z1 = Ey[idx + sz];
    const T ez_x1 = Ez[idx + 1], ez_y1 = Ez[idx + sy];
    const T ex_y1 = Ex[idx + sy], ey_x1 = Ey[idx + 1];
    // x: d_z Ey - d_y Ez ; y: d_x Ez - d_z Ex ; z: d_y Ex - d_x Ey
    const T Kx = p.dt * (p.inv[2] * (ey_z1 - ey) - p.inv[1] * (ez_y1 - ez));
    const T Ky = p.dt * (p.inv[0] * (ez_x1 - ez) - p.inv[2] * (ex_z1 - ex));
    const T Kz = p.dt * (p.inv[1] * (ex_y1 - ex) - p.inv[0] * (ey_x1 - ey));
    // (own, next, prev) = (x,y,z), (y,z,x), (z,x,y)
    cascade_component(p, 0, idx, Kx, sgx, sgy, sgz, gx, gy, gz);
    cascade_component(p, 1, idx, Ky, sgy, sgz, sgx, gx, gy, gz);
    cascade_component(p, 2, idx, Kz, sgz, sgx, sgy, gx, gy, gz);
}

// E half step, cascade: backward differences (idx_curl = -1),
// d(A) = inv * (A[i-1] - A[i]).
template <typename T>
__global__ __launch_bounds__(CBLK) void cascade_e_kernel(CascadeArgs<T> p) {
    int gx, gy, gz;
    if (!cascade_cell(p, gx, gy, gz)) return;
    const int64_t sy = p.sy, sz = p.sz;
    const size_t idx = (size_t)gx + (size_t)gy * sy + (size_t)gz * sz;
    const T sgx = p.sig[0][gx], sgy = p.sig[1][gy], sgz = p.sig[2][gz];
    const T *Hx = p.F[0], *Hy = p.F[1], *Hz = p.F[2];
    const T hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
    const T hx_zm = Hx[idx - sz], hy_zm = Hy[idx - sz];
    const T hz_xm = Hz[idx - 1], hz_ym = Hz[idx - sy];
    const T hx_ym = Hx[idx - sy], hy_xm = Hy[idx - 1];
    // x: d_z Hy - d_y Hz ; y: d_x Hz - d_z Hx ; z: d_y Hx - d_x Hy
    const T Kx = p.dt * (p.inv[2] * (hy_zm - hy) - p.inv[1] * (hz_ym - hz));
    const T Ky = p.dt * (p.inv[0] * (hz_xm - hz) - p.inv[2] * (hx_zm - hx));
    const T Kz = p.dt * (p.inv[1] * (hx_ym - hx) - p.inv[0] * (hy_xm - hy));
    cascade_component(p, 0, idx, Kx, sgx, sgy, sgz, gx, gy, gz);
    cascade_component(p, 1, idx, Ky, sgy, sgz, sgx, gx, gy, gz);
    cascade_component(p, 2, idx, Kz, sgz, sgx, sgy, gx, gy, gz);
}

// ----------------------------------------------------------- inner box
template <typename T>
__device__ __forceinline__ bool inner_lane(const InnerArgs<T> &p, int &x, int &y, int &z0, int &z1,
                                           bool &v0, bool &v1) {
    int L;
    const int ntile = p.tiles[0] * p.tiles[1] * p.tiles[2];
    if (!xcd_block(p.per_xcd, ntile, L)) return false;
    int tx = L % p.tiles[0];
    int r = L / p.tiles[0];
    int ty = r % p.tiles[1];
    int tz = r / p.tiles[1];
    x = p.x0 + tx * (2 * TX) + 2 * threadIdx.x;
    y = p.lo[1] + ty * TY + threadIdx.y;
    if (y > p.hi[1]) return false;
    v0 = x >= p.lo[0] && x <= p.hi[0];
    v1 = x + 1 >= p.lo[0] && x + 1 <= p.hi[0];
    if (!v0 && !v1) return false;
    z0 = p.lo[2] + tz * p.zb;
    z1 = min(z0 + p.zb - 1, p.hi[2]);
    return true;
}

// H half step on the inner box.  Lane owns cells (x, x+1).
template <typename T, bool PERVOX, bool NT>
__global__ __launch_bounds__(256) void inner_h_kernel(InnerArgs<T> p) {
    using V = typename V2<T>::t;
    int x, y, z0, z1;
    bool v0, v1;
    if (!inner_lane(p, x, y, z0, z1, v0, v1)) return;
    const int64_t sy = p.sy, sz = p.sz;
    const T *Ex = p.F[0], *Ey = p.F[1], *Ez = p.F[2];
    T *Hx = p.A[0], *Hy = p.A[1], *Hz = p.A[2];
    const T dt = p.dt, ix = p.inv[0], iy = p.inv[1], iz = p.inv[2];
    size_t idx = (size_t)x + (size_t)y * sy + (size_t)z0 * sz;

    struct Plane {
        V ex1, ey1, ez, ez_y1, ex_y1, hx, hy, hz, mx, my, mz;
        T ez_x2, ey_x2;
    };
    auto load = [&](size_t i) {
        Plane q;
        q.ex1 = ld2(Ex, i + sz);
        q.ey1 = ld2(Ey, i + sz);
        q.ez = ld2(Ez, i);
        q.ez_x2 = Ez[i + 2];
        q.ey_x2 = Ey[i + 2];
        q.ez_y1 = ld2(Ez, i + sy);
        q.ex_y1 = ld2(Ex, i + sy);
        q.hx = ldo<T, NT>((const T *)Hx, i);
        q.hy = ldo<T, NT>((const T *)Hy, i);
        q.hz = ldo<T, NT>((const T *)Hz, i);
        if (PERVOX) {
            q.mx = ldo<T, NT>(p.minv[0], i);
            q.my = ldo<T, NT>(p.minv[1], i);
            q.mz = ldo<T, NT>(p.minv[2], i);
        }
        return q;
    };
    V ex = ld2(Ex, idx), ey = ld2(Ey, idx);
    Plane cur = load(idx);
    for (int z = z0; z <= z1; ++z, idx += sz) {
        Plane nxt;
        if (z < z1) nxt = load(idx + sz);
        const T m0x = PERVOX ? cur.mx.x : p.minv_s, m1x = PERVOX ? cur.mx.y : p.minv_s;
        const T m0y = PERVOX ? cur.my.x : p.minv_s, m1y = PERVOX ? cur.my.y : p.minv_s;
        const T m0z = PERVOX ? cur.mz.x : p.minv_s, m1z = PERVOX ? cur.mz.y : p.minv_s;
        // cell x (element .x): x+1 neighbours are element .y; cell x+1 uses x+2
        const T Kx0 = dt * (iz * (cur.ey1.x - ey.x) - iy * (cur.ez_y1.x - cur.ez.x));
        const T Kx1 = dt * (iz * (cur.ey1.y - ey.y) - iy * (cur.ez_y1.y - cur.ez.y));
        const T Ky0 = dt * (ix * (cur.ez.y - cur.ez.x) - iz * (cur.ex1.x - ex.x));
        const T Ky1 = dt * (ix * (cur.ez_x2 - cur.ez.y) - iz * (cur.ex1.y - ex.y));
        const T Kz0 = dt * (iy * (cur.ex_y1.x - ex.x) - ix * (ey.y - ey.x));
        const T Kz1 = dt * (iy * (cur.ex_y1.y - ex.y) - ix * (cur.ey_x2 - ey.y));
        V nhx, nhy, nhz;
        nhx.x = v0 ? cur.hx.x + m0x * Kx0 : cur.hx.x;
        nhx.y = v1 ? cur.hx.y + m1x * Kx1 : cur.hx.y;
        nhy.x = v0 ? cur.hy.x + m0y * Ky0 : cur.hy.x;
        nhy.y = v1 ? cur.hy.y + m1y * Ky1 : cur.hy.y;
        nhz.x = v0 ? cur.hz.x + m0z * Kz0 : cur.hz.x;
        nhz.y = v1 ? cur.hz.y + m1z * Kz1 : cur.hz.y;
        sto<T, NT>(Hx, idx, nhx);
        sto<T, NT>(Hy, idx, nhy);
        sto<T, NT>(Hz, idx, nhz);
        ex = cur.ex1;
        ey = cur.ey1;
        cur = nxt;
    }
}

// E half step on the inner box.  Lane owns cells (x, x+1).
template <typename T, bool PERVOX, bool NT>
__global__ __launch_bounds__(256) void inner_e_kernel(InnerArgs<T> p) {
    using V = typename V2<T>::t;
    int x, y, z0, z1;
    bool v0, v1;
    if (!inner_lane(p, x, y, z0, z1, v0, v1)) return;
    const int64_t sy = p.sy, sz = p.sz;
    const T *Hx = p.F[0], *Hy = p.F[1], *Hz = p.F[2];
    T *Ex = p.A[0], *Ey = p.A[1], *Ez = p.A[2];
    const T dt = p.dt, ix = p.inv[0], iy = p.inv[1], iz = p.inv[2];
    size_t idx = (size_t)x + (size_t)y * sy + (size_t)z0 * sz;

    struct Plane {
        V hx, hy, hz, hz_ym, hx_ym, ex, ey, ez, mx, my, mz;
        T hz_xm, hy_xm;
    };
    auto load = [&](size_t i) {
        Plane q;
        q.hx = ld2(Hx, i);
        q.hy = ld2(Hy, i);
        q.hz = ld2(Hz, i);
        q.hz_xm = Hz[i - 1];
        q.hy_xm = Hy[i - 1];
        q.hz_ym = ld2(Hz, i - sy);
        q.hx_ym = ld2(Hx, i - sy);
        q.ex = ldo<T, NT>((const T *)Ex, i);
        q.ey = ldo<T, NT>((const T *)Ey, i);
        q.ez = ldo<T, NT>((const T *)Ez, i);
        if (PERVOX) {
            q.mx = ldo<T, NT>(p.minv[0], i);
            q.my = ldo<T, NT>(p.minv[1], i);
            q.mz = ldo<T, NT>(p.minv[2], i);
        }
        return q;
    };
    V hxm = ld2(Hx, idx - sz), hym = ld2(Hy, idx - sz);
    Plane cur = load(idx);
    for (int z = z0; z <= z1; ++z, idx += sz) {
        Plane nxt;
        if (z < z1) nxt = load(idx + sz);
        const T m0x = PERVOX ? cur.mx.x : p.minv_s, m1x = PERVOX ? cur.mx.y : p.minv_s;
        const T m0y = PERVOX ? cur.my.x : p.minv_s, m1y = PERVOX ? cur.my.y : p.minv_s;
        const T m0z = PERVOX ? cur.mz.x : p.minv_s, m1z = PERVOX ? cur.mz.y : p.minv_s;
        // cell x (element .x): x-1 neighbour is the scalar; cell x+1 uses element .x
        const T Kx0 = dt * (iz * (hym.x - cur.hy.x) - iy * (cur.hz_ym.x - cur.hz.x));
        const T Kx1 = dt * (iz * (hym.y - cur.hy.y) - iy * (cur.hz_ym.y - cur.hz.y));
        const T Ky0 = dt * (ix * (cur.hz_xm - cur.hz.x) - iz * (hxm.x - cur.hx.x));
        const T Ky1 = dt * (ix * (cur.hz.x - cur.hz.y) - iz * (hxm.y - cur.hx.y));
        const T Kz0 = dt * (iy * (cur.hx_ym.x - cur.hx.x) - ix * (cur.hy_xm - cur.hy.x));
        const T Kz1 = dt * (iy * (cur.hx_ym.y - cur.hx.y) - ix * (cur.hy.x - cur.hy.y));
        V nex, ney, nez;
        nex.x = v0 ? cur.ex.x + m0x * Kx0 : cur.ex.x;
        nex.y = v1 ? cur.ex.y + m1x * Kx1 : cur.ex.y;
        ney.x = v0 ? cur.ey.x + m0y * Ky0 : cur.ey.x;
        ney.y = v1 ? cur.ey.y + m1y * Ky1 : cur.ey.y;
        nez.x = v0 ? cur.ez.x + m0z * Kz0 : cur.ez.x;
        nez.y = v1 ? cur.ez.y + m1z * Kz1 : cur.ez.y;
        sto<T, NT>(Ex, idx, nex);
        sto<T, NT>(Ey, idx, ney);
        sto<T, NT>(Ez, idx, nez);
        hxm = cur.hx;
        hym = cur.hy;
        cur = nxt;
    }
}

// ------------------------------------------------------- FAST-mode PML
// D/B-eliminated cascade.  The reference keeps T (= D or B) and sets
// A = m (T + S) where the own-axis sigma is zero, W = m (T + S) otherwise
// (src/Kernels/Helpers.jl:323-340).  Both relations let the T stage run on
// A (or W) directly:
//   U' = ((1 - s_n) U + K) / (1 + s_n),  in = U' - U        (s_n != 0)
//   s_o == 0:  A' = ((1 - s_p) A + m in) / (1 + s_p)
//   s_o != 0:  W' = ((1 - s_p) W + m in) / (1 + s_p),
//              A' = A + (1 + s_o) W' - (1 - s_o) W
// Source terms are added as the values are stored (PmlArgs::src, up to
// MAX_SRC per half step) or afterwards by source_fix_kernel.  A lane owns the x
// pair (x, x+1) like the inner kernels; slab x ranges are widened to whole
// pairs and the extra element is written back unchanged.  With complex
// fields (TWO kernels) a lane updates the real part and then the imaginary
// part of its pair: the material bytes, sigma values and indexing are read
// and computed once for both.
// A source box whose FAST corrections for a block of steps were computed
// ahead of the updates (source_dw_kernel): the block step's row holds the
// boxes one after another, each dense with x fastest, from `off` on.
struct PreBox {
    int s[3], d[3];
    int64_t off;
};

struct PSlab {
    int lo[3];      // first cell: lo[0] odd
    int n[3];       // extent: n[0] even
    int xlo, xhi;   // cells actually updated along x
    int first;      // first logical block
};

// One part's arrays: A updated from the curl of F, its U and W stages; with
// a periodic x axis folded into the kernel (perx), the x-neighbour across
// the boundary is read from the opposite interior plane as the wrap would
// have filled the ghost: wa F + wb Fo in float64 (Fo the other part of a
// complex field or null; wa, wb the Bloch phase for this part).
template <typename T>
struct PmlPart {
    T *A[3];
    const T *F[3];
    T *U[3];
    T *W[3];
    const T *Fo[3];
    double wa, wb;
};

// A small E monitor whose DFT the E update accumulates (real fields, float64
// accumulators): its columns in a row of the step weight table are [due,
// then dt cos and dt sin of f_k theta per frequency].
struct DfteMon {
    double *acc;
    int comp;
    int s[3], d[3];
    int nf, woff;
    int64_t ncell;
};

template <typename T>
struct PmlArgs {
    PmlPart<T> f[2];  // the part updated; [1]: the imaginary part (TWO kernels)
    const T *sig[3];
    const T *minv[3];
    T minv_s;
    T dt;
    T inv[3];
    int64_t sy, sz;
    int nslab;
    PSlab slab[MAX_SLAB];
    int nblock;
    int per_xcd;
    // conductivity (COND kernels): per-voxel sigma on the padded grid and the
    // C stage arrays, touched only inside the box clo..chi where some sigma
    // of the group is nonzero
    const T *msig[3];
    T *C[3];
    int clo[3], chi[3];
    // per-voxel material as one byte per cell into a table of at most 256
    // values (LUT kernels): the same numbers as minv, a quarter of the bytes
    const uint8_t *mid[3];
    const T *lut;
    // periodic x folded into the kernel (see PmlPart); nxg = N along x
    int perx;
    int nxg;
    // FAST-mode sources of this half step applied to the part f[0] as the
    // update stores its values: source_fix_kernel's correction, in
    // registration order, on the cells inside each source's box (nsrc = 0:
    // none, or the host launches source_fix_kernel instead)
    int nsrc;
    FixSrc<T> src[MAX_SRC];
    const int64_t *clk;
    int koff;
    double cutoff;
    // or: the half step's sources, all on component pre_comp with pairwise
    // disjoint boxes, as corrections computed for the block of steps ahead of
    // time: the value at box cell i of block step koff is
    // pre_dw[koff * pre_cells + pre[q].off + i], loaded with the lane's other
    // values and added to the updated value (and W) before the store
    int pre_n, pre_comp;
    PreBox pre[MAX_SRC];
    const T *pre_dw;
    int64_t pre_cells;
    // SNAP kernels (E update of the adjoint's re-run): each component's
    // stored value (W where the own-axis sigma is non-zero, else A) on the
    // cells of box snap_lo[a] .. + snap_n[a] also goes to snap[a], dense x
    // fastest, as box_copy_kernel would copy it after the step (null: none)
    T *snap[3];
    int snap_lo[3][3], snap_n[3][3];
    // DFTE kernels: the DFT update of these E monitors at this step, with
    // weights from dftw (this step's row of the step weight table), as
    // dft_kernel would apply it after the half step
    int ndfte;
    DfteMon dfte[MAX_DFTE];
    const double *dftw;
};

// the ghost value wrap_kernel would write from interior element i
template <typename T>
__device__ __forceinline__ T xwrap_val(const PmlPart<T> &f, int a, size_t i) {
    double v = f.wa * (double)f.F[a][i];
    if (f.Fo[a]) v += f.wb * (double)f.Fo[a][i];
    return (T)v;
}

template <typename T>
__device__ __forceinline__ bool pml_lane(const PmlArgs<T> &p, int &x, int &y, int &z, bool &v0, bool &v1) {
    int L;
    if (!xcd_block(p.per_xcd, p.nblock, L)) return false;
    int s = 0;
    while (s + 1 < p.nslab && L >= p.slab[s + 1].first) ++s;
    const PSlab &b = p.slab[s];
    const int npx = b.n[0] >> 1;
    const int64_t c = (int64_t)(L - b.first) * CBLK + threadIdx.x;
    if (c >= (int64_t)npx * b.n[1] * b.n[2]) return false;
    x = b.lo[0] + 2 * (int)(c % npx);
    const int64_t r = c / npx;
    y = b.lo[1] + (int)(r % b.n[1]);
    z = b.lo[2] + (int)(r / b.n[1]);
    v0 = x >= b.xlo && x <= b.xhi;
    v1 = x + 1 >= b.xlo && x + 1 <= b.xhi;
    return v0 || v1;
}

// State of one component of the pair, loaded before any store of the lane
// (the memory counter is in order: a load issued after a store waits for it).
template <typename T>
struct PmlComp {
    typename V2<T>::t m, ao, uo, wo;
    typename V2<T>::t co, sd;  // C stage: old C, sigma~ = 0.5 (dt sigma)
    bool nu, nw;  // U stage (next-axis sigma) / W stage (own-axis sigma) present
    bool nc;      // C stage (lane inside the conductivity box)
};

template <typename T, bool PERVOX, bool LUT>
__device__ __forceinline__ void pml_load_m(const PmlArgs<T> &p, int a, size_t idx, PmlComp<T> &c) {
    if (LUT) {
        const unsigned short q = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p.mid[a] + idx));
        c.m.x = p.lut[q & 0xff];
        c.m.y = p.lut[q >> 8];
    } else if (PERVOX) {
        c.m = ldo<T, true>(p.minv[a], idx);
    } else {
        c.m.x = p.minv_s;
        c.m.y = p.minv_s;
    }
}

template <typename T>
__device__ __forceinline__ void pml_load_uw(const PmlPart<T> &f, int a, size_t idx, bool nu, bool nw,
                                            PmlComp<T> &c) {
    c.nu = nu;
    c.nw = nw;
    if (nu) c.uo = ldo<T, true>((const T *)f.U[a], idx);
    if (nw) c.wo = ldo<T, true>((const T *)f.W[a], idx);
}

// The precomputed corrections of the lane's pair (x, x+1) in row y, plane z
// (at most one box holds a cell).
template <typename T>
__device__ __forceinline__ void pml_pre_load(const PmlArgs<T> &p, int x, int y, int z, bool v0, bool v1, T &dw0,
                                             T &dw1, bool &h0, bool &h1) {
    const T *row = p.pre_dw + (int64_t)p.koff * p.pre_cells;
    for (int q = 0; q < p.pre_n; ++q) {
        const PreBox &b = p.pre[q];
        if (y < b.s[1] || y >= b.s[1] + b.d[1] || z < b.s[2] || z >= b.s[2] + b.d[2]) continue;
        const int64_t r = b.off + (int64_t)b.d[0] * ((int64_t)(y - b.s[1]) + (int64_t)b.d[1] * (z - b.s[2]));
        const int i0 = x - b.s[0];
        if (v0 && i0 >= 0 && i0 < b.d[0]) {
            dw0 = row[r + i0];
            h0 = true;
        }
        if (v1 && i0 + 1 >= 0 && i0 + 1 < b.d[0]) {
            dw1 = row[r + i0 + 1];
            h1 = true;
        }
    }
}

// The FAST source correction at the lane's pair (x, x+1) in row y, plane z,
// once the update has stored its values: source by source in registration
// order, the same loads and operations as source_fix_kernel (the lane reads
// back what it stored), so the result is that of the separate launch.
template <typename T>
__device__ __forceinline__ void pml_sources(const PmlArgs<T> &p, size_t idx, int x, int y, int z, bool v0,
                                            bool v1) {
    for (int q = 0; q < p.nsrc; ++q) {
        const FixSrc<T> &o = p.src[q];
        if (y < o.s[1] || y >= o.s[1] + o.d[1] || z < o.s[2] || z >= o.s[2] + o.d[2]) continue;
        for (int e = 0; e < 2; ++e) {
            const int g[3] = {x + e, y, z};
            if (!(e ? v1 : v0) || g[0] < o.s[0] || g[0] >= o.s[0] + o.d[0]) continue;
            const int64_t ts = *p.clk + p.koff;
            const bool now = sources_on(ts, p.dt, p.cutoff);
            const bool before = ts >= 1 && sources_on(ts - 1, p.dt, p.cutoff);
            if (!now && !before) continue;
            const size_t i = idx + (size_t)e;
            const size_t ci = (size_t)(g[0] - o.s[0]) +
                              (size_t)o.d[0] * ((size_t)(y - o.s[1]) + (size_t)o.d[1] * (size_t)(z - o.s[2]));
            const T ar = o.amp[2 * ci], ai = o.amp[2 * ci + 1];
            const T m = o.minv ? o.minv[i] : p.minv_s;
            const T so = o.sig_own[g[o.own]], sp = o.sig_prev[g[o.prev]];
            const T ao = o.A[i];
            const T wo = so == T(0) ? T(0) : o.W[i];
            T fr = 0, fi = 0, pr = 0, pi = 0;
            if (now) wave_value(o.wave, o.nwave, ts, fr, fi);
            if (before) wave_value(o.wave, o.nwave, ts - 1, pr, pi);
            const T vn = (fr * ar) - (fi * ai);
            const T vo = (pr * ar) - (pi * ai);
            const T dw = m * (vn - ((T(1) - sp) / (T(1) + sp)) * vo);
            if (so == T(0)) {
                o.A[i] = ao + dw;
            } else {
                o.W[i] = wo + dw;
                o.A[i] = ao + (T(1) + so) * dw;
            }
        }
    }
}

// Update and store one component.  so/sn/sp: own/next/prev sigma per element.
// dw0/dw1 with h0/h1: a precomputed source correction of the element, added
// as source_fix_kernel would add it after the store.
template <typename T>
__device__ __forceinline__ void pml_update(const PmlArgs<T> &p, const PmlPart<T> &f, int a, size_t idx,
                                           const PmlComp<T> &c, T K0, T K1, T so0, T so1, T sn0, T sn1, T sp0,
                                           T sp1, bool v0, bool v1, T dw0 = T(0), T dw1 = T(0), bool h0 = false,
                                           bool h1 = false, typename V2<T>::t *kept = nullptr,
                                           typename V2<T>::t *stored = nullptr) {
    using V = typename V2<T>::t;
    T in0 = K0, in1 = K1;
    if (c.nc) {
        // C stage (src/Kernels/Helpers.jl:58-63): dC/dt + sigma_D C = K
        V cn = c.co;
        if (c.sd.x != T(0) && v0) {
            cn.x = ((T(1) - c.sd.x) * c.co.x + K0) / (T(1) + c.sd.x);
            in0 = cn.x - c.co.x;
        }
        if (c.sd.y != T(0) && v1) {
            cn.y = ((T(1) - c.sd.y) * c.co.y + K1) / (T(1) + c.sd.y);
            in1 = cn.y - c.co.y;
        }
        sto<T, true>(p.C[a], idx, cn);
    }
    if (c.nu) {
        V un = c.uo;
        if (sn0 != T(0) && v0) {
            un.x = ((T(1) - sn0) * c.uo.x + in0) / (T(1) + sn0);
            in0 = un.x - c.uo.x;
        }
        if (sn1 != T(0) && v1) {
            un.y = ((T(1) - sn1) * c.uo.y + in1) / (T(1) + sn1);
            in1 = un.y - c.uo.y;
        }
        sto<T, true>(f.U[a], idx, un);
    }
    V an = c.ao;
    V wn = c.wo;  // loaded and stored with the W stage only
    if (c.nw) {
        if (v0) {
            if (so0 != T(0)) {
                wn.x = ((T(1) - sp0) * c.wo.x + c.m.x * in0) / (T(1) + sp0);
                an.x = (c.ao.x + (T(1) + so0) * wn.x) - (T(1) - so0) * c.wo.x;
            } else {
                an.x = ((T(1) - sp0) * c.ao.x + c.m.x * in0) / (T(1) + sp0);
            }
        }
        if (v1) {
            if (so1 != T(0)) {
                wn.y = ((T(1) - sp1) * c.wo.y + c.m.y * in1) / (T(1) + sp1);
                an.y = (c.ao.y + (T(1) + so1) * wn.y) - (T(1) - so1) * c.wo.y;
            } else {
                an.y = ((T(1) - sp1) * c.ao.y + c.m.y * in1) / (T(1) + sp1);
            }
        }
    } else {
        if (v0) an.x = ((T(1) - sp0) * c.ao.x + c.m.x * in0) / (T(1) + sp0);
        if (v1) an.y = ((T(1) - sp1) * c.ao.y + c.m.y * in1) / (T(1) + sp1);
    }
    // (a non-zero own-axis sigma implies the W stage)
    if (h0) {
        if (so0 == T(0)) {
            an.x = an.x + dw0;
        } else {
            wn.x = wn.x + dw0;
            an.x = an.x + (T(1) + so0) * dw0;
        }
    }
    if (h1) {
        if (so1 == T(0)) {
            an.y = an.y + dw1;
        } else {
            wn.y = wn.y + dw1;
            an.y = an.y + (T(1) + so1) * dw1;
        }
    }
    if (c.nw) sto<T, true>(f.W[a], idx, wn);
    sto<T, true>(f.A[a], idx, an);
    if (kept) {
        kept->x = so0 != T(0) ? wn.x : an.x;
        kept->y = so1 != T(0) ? wn.y : an.y;
    }
    if (stored) *stored = an;
}

// The DFT update of the E monitors of component a at the lane's pair, from
// the values the update stored (DFTE kernels): the same operations on the
// same values as dft_kernel (acc += w_k * value, w_k = dt exp(i f_k theta)).
template <typename T>
__device__ __forceinline__ void pml_dft(const PmlArgs<T> &p, int a, int x, int y, int z, bool v0, bool v1,
                                        typename V2<T>::t v) {
    for (int q = 0; q < p.ndfte; ++q) {
        const DfteMon &m = p.dfte[q];
        if (m.comp != a || y < m.s[1] || y >= m.s[1] + m.d[1] || z < m.s[2] || z >= m.s[2] + m.d[2]) continue;
        const double *w = p.dftw + m.woff;
        for (int e = 0; e < 2; ++e) {
            const int i = x + e - m.s[0];
            if (!(e ? v1 : v0) || i < 0 || i >= m.d[0]) continue;
            if (w[0] == 0.0) continue;
            const double u = (double)(e ? v.y : v.x);
            double *acc = m.acc + 2 * ((int64_t)i + (int64_t)m.d[0] * ((int64_t)(y - m.s[1]) +
                                                                      (int64_t)m.d[1] * (z - m.s[2])));
            for (int k = 0; k < m.nf; ++k, acc += 2 * m.ncell) {
                const double re = acc[0] + w[1 + 2 * k] * u;
                const double im = acc[1] + w[2 + 2 * k] * u;
                acc[0] = re;
                acc[1] = im;
            }
        }
    }
}

// The design-box snapshot of component a at the lane's pair (SNAP kernels).
template <typename T>
__device__ __forceinline__ void pml_snap(const PmlArgs<T> &p, int a, int x, int y, int z, bool v0, bool v1,
                                         typename V2<T>::t v) {
    T *o = p.snap[a];
    if (!o) return;
    const int *lo = p.snap_lo[a], *n = p.snap_n[a];
    if (y < lo[1] || y >= lo[1] + n[1] || z < lo[2] || z >= lo[2] + n[2]) return;
    const int64_t r = (int64_t)n[0] * ((int64_t)(y - lo[1]) + (int64_t)n[1] * (z - lo[2]));
    const int i0 = x - lo[0];
    if (v0 && i0 >= 0 && i0 < n[0]) o[r + i0] = v.x;
    if (v1 && i0 + 1 >= 0 && i0 + 1 < n[0]) o[r + i0 + 1] = v.y;
}

// Loads of one part's state for the three components (the material values
// m are loaded once per lane by pml_load_m):
// (own, next, prev) = (x,y,z), (y,z,x), (z,x,y).
template <typename T, bool COND>
__device__ __forceinline__ void pml_load_all(const PmlArgs<T> &p, const PmlPart<T> &f, size_t idx, T sx0, T sx1,
                                             T sgy, T sgz, int x, int y, int z, PmlComp<T> c[3]) {
    c[0].ao = ldo<T, true>((const T *)f.A[0], idx);
    c[1].ao = ldo<T, true>((const T *)f.A[1], idx);
    c[2].ao = ldo<T, true>((const T *)f.A[2], idx);
    const bool px = sx0 != T(0) || sx1 != T(0), py = sgy != T(0), pz = sgz != T(0);
    pml_load_uw(f, 0, idx, py, px, c[0]);
    pml_load_uw(f, 1, idx, pz, py, c[1]);
    pml_load_uw(f, 2, idx, px, pz, c[2]);
    const bool inbox = COND && x + 1 >= p.clo[0] && x <= p.chi[0] && y >= p.clo[1] && y <= p.chi[1] &&
                       z >= p.clo[2] && z <= p.chi[2];
    for (int a = 0; a < 3; ++a) {
        c[a].nc = inbox && p.C[a] != nullptr;
        if (c[a].nc) {
            c[a].co = ldo<T, true>((const T *)p.C[a], idx);
            const auto sg = ldo<T, true>(p.msig[a], idx);
            c[a].sd.x = T(0.5) * (p.dt * sg.x);
            c[a].sd.y = T(0.5) * (p.dt * sg.y);
        }
    }
}

// One part of the H update at the lane's pair (x, x+1), row y, plane z.
template <typename T, bool COND>
__device__ __forceinline__ void pml_h_part(const PmlArgs<T> &p, const PmlPart<T> &f, size_t idx, int x, int y,
                                           int z, T sx0, T sx1, T sgy, T sgz, bool v0, bool v1, PmlComp<T> c[3],
                                           T dw0 = T(0), T dw1 = T(0), bool h0 = false, bool h1 = false) {
    const int64_t sy = p.sy, sz = p.sz;
    const T *Ex = f.F[0], *Ey = f.F[1], *Ez = f.F[2];
    const auto ex = ld2(Ex, idx);
    auto ey = ld2(Ey, idx), ez = ld2(Ez, idx);
    const auto ex1 = ld2(Ex, idx + sz), ey1 = ld2(Ey, idx + sz);
    const auto ez_y1 = ld2(Ez, idx + sy), ex_y1 = ld2(Ex, idx + sy);
    T ez_x2 = Ez[idx + 2], ey_x2 = Ey[idx + 2];
    if (p.perx) {
        // E at x = N + 1 is E at x = 1 times the Bloch phase
        if (x + 1 == p.nxg) {
            const size_t i1 = idx + 2 - (size_t)p.nxg;
            ez_x2 = xwrap_val(f, 2, i1);
            ey_x2 = xwrap_val(f, 1, i1);
        } else if (x == p.nxg) {
            const size_t i1 = idx + 1 - (size_t)p.nxg;
            ez.y = xwrap_val(f, 2, i1);
            ey.y = xwrap_val(f, 1, i1);
        }
    }
    pml_load_all<T, COND>(p, f, idx, sx0, sx1, sgy, sgz, x, y, z, c);
    const T dt = p.dt, ix = p.inv[0], iy = p.inv[1], iz = p.inv[2];
    const T Kx0 = dt * (iz * (ey1.x - ey.x) - iy * (ez_y1.x - ez.x));
    const T Kx1 = dt * (iz * (ey1.y - ey.y) - iy * (ez_y1.y - ez.y));
    const T Ky0 = dt * (ix * (ez.y - ez.x) - iz * (ex1.x - ex.x));
    const T Ky1 = dt * (ix * (ez_x2 - ez.y) - iz * (ex1.y - ex.y));
    const T Kz0 = dt * (iy * (ex_y1.x - ex.x) - ix * (ey.y - ey.x));
    const T Kz1 = dt * (iy * (ex_y1.y - ex.y) - ix * (ey_x2 - ey.y));
    const int pc = p.pre_comp;
    pml_update(p, f, 0, idx, c[0], Kx0, Kx1, sx0, sx1, sgy, sgy, sgz, sgz, v0, v1, dw0, dw1, h0 && pc == 0,
               h1 && pc == 0);
    pml_update(p, f, 1, idx, c[1], Ky0, Ky1, sgy, sgy, sgz, sgz, sx0, sx1, v0, v1, dw0, dw1, h0 && pc == 1,
               h1 && pc == 1);
    pml_update(p, f, 2, idx, c[2], Kz0, Kz1, sgz, sgz, sx0, sx1, sgy, sgy, v0, v1, dw0, dw1, h0 && pc == 2,
               h1 && pc == 2);
}

// One part of the E update at the lane's pair.
template <typename T, bool COND, bool SNAP = false, bool DFTE = false>
__device__ __forceinline__ void pml_e_part(const PmlArgs<T> &p, const PmlPart<T> &f, size_t idx, int x, int y,
                                           int z, T sx0, T sx1, T sgy, T sgz, bool v0, bool v1, PmlComp<T> c[3],
                                           T dw0 = T(0), T dw1 = T(0), bool h0 = false, bool h1 = false) {
    const int64_t sy = p.sy, sz = p.sz;
    const T *Hx = f.F[0], *Hy = f.F[1], *Hz = f.F[2];
    const auto hx = ld2(Hx, idx), hy = ld2(Hy, idx), hz = ld2(Hz, idx);
    const auto hxm = ld2(Hx, idx - sz), hym = ld2(Hy, idx - sz);
    const auto hz_ym = ld2(Hz, idx - sy), hx_ym = ld2(Hx, idx - sy);
    T hz_xm = Hz[idx - 1], hy_xm = Hy[idx - 1];
    if (p.perx && x == 1) {
        // H at x = 0 is H at x = N times the Bloch phase
        const size_t i0 = idx - 1 + (size_t)p.nxg;
        hz_xm = xwrap_val(f, 2, i0);
        hy_xm = xwrap_val(f, 1, i0);
    }
    pml_load_all<T, COND>(p, f, idx, sx0, sx1, sgy, sgz, x, y, z, c);
    const T dt = p.dt, ix = p.inv[0], iy = p.inv[1], iz = p.inv[2];
    const T Kx0 = dt * (iz * (hym.x - hy.x) - iy * (hz_ym.x - hz.x));
    const T Kx1 = dt * (iz * (hym.y - hy.y) - iy * (hz_ym.y - hz.y));
    const T Ky0 = dt * (ix * (hz_xm - hz.x) - iz * (hxm.x - hx.x));
    const T Ky1 = dt * (ix * (hz.x - hz.y) - iz * (hxm.y - hx.y));
    const T Kz0 = dt * (iy * (hx_ym.x - hx.x) - ix * (hy_xm - hy.x));
    const T Kz1 = dt * (iy * (hx_ym.y - hx.y) - ix * (hy.x - hy.y));
    const int pc = p.pre_comp;
    typename V2<T>::t kept[3], stored[3];
    pml_update(p, f, 0, idx, c[0], Kx0, Kx1, sx0, sx1, sgy, sgy, sgz, sgz, v0, v1, dw0, dw1, h0 && pc == 0,
               h1 && pc == 0, SNAP ? &kept[0] : nullptr, DFTE ? &stored[0] : nullptr);
    pml_update(p, f, 1, idx, c[1], Ky0, Ky1, sgy, sgy, sgz, sgz, sx0, sx1, v0, v1, dw0, dw1, h0 && pc == 1,
               h1 && pc == 1, SNAP ? &kept[1] : nullptr, DFTE ? &stored[1] : nullptr);
    pml_update(p, f, 2, idx, c[2], Kz0, Kz1, sgz, sgz, sx0, sx1, sgy, sgy, v0, v1, dw0, dw1, h0 && pc == 2,
               h1 && pc == 2, SNAP ? &kept[2] : nullptr, DFTE ? &stored[2] : nullptr);
    if (SNAP)
        for (int a = 0; a < 3; ++a) pml_snap(p, a, x, y, z, v0, v1, kept[a]);
    if (DFTE)
        for (int a = 0; a < 3; ++a) pml_dft(p, a, x, y, z, v0, v1, stored[a]);
}

// TWO: complex fields, p.f[0] the real and p.f[1] the imaginary part (no
// conductivity: its C stage is per part).  PRE: the launch adds precomputed
// source corrections (real fields without conductivity; a separate variant so
// that the others keep their register count).  SNAP (E update, real fields
// without conductivity): the launch also writes the design-box snapshot;
// DFTE (likewise): it also applies the DFT update of small E monitors.
template <typename T, bool PERVOX, bool COND, bool LUT = false, bool TWO = false, bool PRE = false>
__global__ __launch_bounds__(CBLK) void pml_h_kernel(PmlArgs<T> p) {
    static_assert(!(COND && TWO), "the conductive C stage runs one part per launch");
    static_assert(!(PRE && (COND || TWO)), "precomputed sources: real fields without conductivity");
    int x, y, z;
    bool v0, v1;
    if (!pml_lane(p, x, y, z, v0, v1)) return;
    const size_t idx = (size_t)x + (size_t)y * p.sy + (size_t)z * p.sz;
    const T sx0 = p.sig[0][x], sx1 = p.sig[0][x + 1], sgy = p.sig[1][y], sgz = p.sig[2][z];
    T dw0 = T(0), dw1 = T(0);
    bool h0 = false, h1 = false;
    if (PRE) pml_pre_load(p, x, y, z, v0, v1, dw0, dw1, h0, h1);
    PmlComp<T> c[3];
    pml_load_m<T, PERVOX, LUT>(p, 0, idx, c[0]);
    pml_load_m<T, PERVOX, LUT>(p, 1, idx, c[1]);
    pml_load_m<T, PERVOX, LUT>(p, 2, idx, c[2]);
    pml_h_part<T, COND>(p, p.f[0], idx, x, y, z, sx0, sx1, sgy, sgz, v0, v1, c, dw0, dw1, h0, h1);
    if (TWO) pml_h_part<T, false>(p, p.f[1], idx, x, y, z, sx0, sx1, sgy, sgz, v0, v1, c);
    if (p.nsrc > 0) pml_sources(p, idx, x, y, z, v0, v1);
}

template <typename T, bool PERVOX, bool COND, bool LUT = false, bool TWO = false, bool PRE = false,
          bool SNAP = false, bool DFTE = false>
__global__ __launch_bounds__(CBLK) void pml_e_kernel(PmlArgs<T> p) {
    static_assert(!(COND && TWO), "the conductive C stage runs one part per launch");
    static_assert(!(PRE && (COND || TWO)), "precomputed sources: real fields without conductivity");
    static_assert(!((SNAP || DFTE) && (COND || TWO)), "snapshots, DFTs: real fields without conductivity");
    int x, y, z;
    bool v0, v1;
    if (!pml_lane(p, x, y, z, v0, v1)) return;
    const size_t idx = (size_t)x + (size_t)y * p.sy + (size_t)z * p.sz;
    const T sx0 = p.sig[0][x], sx1 = p.sig[0][x + 1], sgy = p.sig[1][y], sgz = p.sig[2][z];
    T dw0 = T(0), dw1 = T(0);
    bool h0 = false, h1 = false;
    if (PRE) pml_pre_load(p, x, y, z, v0, v1, dw0, dw1, h0, h1);
    PmlComp<T> c[3];
    pml_load_m<T, PERVOX, LUT>(p, 0, idx, c[0]);
    pml_load_m<T, PERVOX, LUT>(p, 1, idx, c[1]);
    pml_load_m<T, PERVOX, LUT>(p, 2, idx, c[2]);
    pml_e_part<T, COND, SNAP, DFTE>(p, p.f[0], idx, x, y, z, sx0, sx1, sgy, sgz, v0, v1, c, dw0, dw1, h0, h1);
    if (TWO) pml_e_part<T, false>(p, p.f[1], idx, x, y, z, sx0, sx1, sgy, sgz, v0, v1, c);
    if (p.nsrc > 0) pml_sources(p, idx, x, y, z, v0, v1);
}

// FAST mode sources.  With T eliminated the reference's E = m (T + S_n)
// (or W = m (T + S_n)) becomes, for one source added after the update,
//   dW = m (S_n - (1 - s_p) / (1 + s_p) S_{n-1});
//   s_o == 0: A += dW;   s_o != 0: W += dW, A += (1 + s_o) dW.
// In the inner box (all sigma zero) this is A += m (S_n - S_{n-1}).
// S = Re(f * amp).  Sources go in one launch as consecutive block ranges,
// in registration order, as long as no two of them in the launch write the
// same cell of one component; the host splits sources that do into
// successive launches, so every overlapped cell takes them in turn.

template <typename T>
struct FixArgs {
    int nsrc;
    FixSrc<T> src[MAX_SRC];
    T minv_s;
    T dt;
    double cutoff;
    const int64_t *clk;
    int koff;
    int n0, n1, n2;
    int64_t sy, sz;
};

template <typename T>
__global__ __launch_bounds__(256) void source_fix_kernel(FixArgs<T> p) {
    int q = 0;
    const int b = blockIdx.x;
    while (q + 1 < p.nsrc && b >= p.src[q + 1].first) ++q;
    const FixSrc<T> &o = p.src[q];
    const int64_t n = (int64_t)o.d[0] * o.d[1] * o.d[2];
    const int64_t c = (int64_t)(b - o.first) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int ix = (int)(c % o.d[0]);
    const int64_t r = c / o.d[0];
    const int iy = (int)(r % o.d[1]), iz = (int)(r / o.d[1]);
    const int g[3] = {o.s[0] + ix, o.s[1] + iy, o.s[2] + iz};
    // only updated cells consume sources (grid index 1..N)
    if (g[0] < 1 || g[0] > p.n0 || g[1] < 1 || g[1] > p.n1 || g[2] < 1 || g[2] > p.n2) return;
    // everything that does not depend on the timestep is loaded together
    // with the timestep, so the lane waits for memory twice (timestep, then
    // the waveform tables) rather than once per dependent load
    const size_t idx = (size_t)g[0] + (size_t)g[1] * p.sy + (size_t)g[2] * p.sz;
    const T ar = o.amp[2 * c], ai = o.amp[2 * c + 1];
    const T m = o.minv ? o.minv[idx] : p.minv_s;
    const T so = o.sig_own[g[o.own]], sp = o.sig_prev[g[o.prev]];
    const T ao = o.A[idx];
    const T wo = so == T(0) ? T(0) : o.W[idx];
    const int64_t ts = *p.clk + p.koff;
    const bool now = sources_on(ts, p.dt, p.cutoff);
    const bool before = ts >= 1 && sources_on(ts - 1, p.dt, p.cutoff);
    if (!now && !before) return;
    T fr = 0, fi = 0, pr = 0, pi = 0;
    if (now) wave_value(o.wave, o.nwave, ts, fr, fi);
    if (before) wave_value(o.wave, o.nwave, ts - 1, pr, pi);
    const T vn = (fr * ar) - (fi * ai);
    const T vo = (pr * ar) - (pi * ai);
    const T dw = m * (vn - ((T(1) - sp) / (T(1) + sp)) * vo);
    if (so == T(0)) {
        o.A[idx] = ao + dw;
    } else {
        o.W[idx] = wo + dw;
        o.A[idx] = ao + (T(1) + so) * dw;
    }
}

// The FAST corrections of a half step's sources for the nsteps steps of a
// block, ahead of the updates: row k (block step k, timestep = device clock +
// k) holds the boxes one after another; each value is the dw source_fix_kernel
// would add at that cell and step, or -0 (the exact identity of an addition)
// where the sources are off.  Cells outside 1..N are never read.
template <typename T>
struct PreSrcArgs {
    int n;
    PreBox box[MAX_SRC];
    const T *amp[MAX_SRC];
    const double *wave[MAX_SRC];
    int64_t nwave[MAX_SRC];
    int prev;
    const T *minv;        // per-voxel or null (minv_s)
    T minv_s;
    const T *sig_prev;
    T dt;
    double cutoff;
    const int64_t *clk;
    int nsteps;
    int64_t cells;        // per row
    int n0, n1, n2;
    int64_t sy, sz;
    T *dw;                // nsteps rows
};

template <typename T>
__global__ __launch_bounds__(256) void source_dw_kernel(PreSrcArgs<T> p) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)p.nsteps * p.cells) return;
    const int k = (int)(t / p.cells);
    const int64_t ci = t - (int64_t)k * p.cells;
    int q = 0;
    while (q + 1 < p.n && ci >= p.box[q + 1].off) ++q;
    const PreBox &b = p.box[q];
    const int64_t li = ci - b.off;
    const int ix = (int)(li % b.d[0]);
    const int64_t r = li / b.d[0];
    const int iy = (int)(r % b.d[1]), iz = (int)(r / b.d[1]);
    const int g[3] = {b.s[0] + ix, b.s[1] + iy, b.s[2] + iz};
    T out = -T(0);
    if (g[0] >= 1 && g[0] <= p.n0 && g[1] >= 1 && g[1] <= p.n1 && g[2] >= 1 && g[2] <= p.n2) {
        const size_t idx = (size_t)g[0] + (size_t)g[1] * p.sy + (size_t)g[2] * p.sz;
        const T ar = p.amp[q][2 * li], ai = p.amp[q][2 * li + 1];
        const T m = p.minv ? p.minv[idx] : p.minv_s;
        const T sp = p.sig_prev[g[p.prev]];
        const int64_t ts = *p.clk + k;
        const bool now = sources_on(ts, p.dt, p.cutoff);
        const bool before = ts >= 1 && sources_on(ts - 1, p.dt, p.cutoff);
        if (now || before) {
            T fr = 0, fi = 0, pr = 0, pi = 0;
            if (now) wave_value(p.wave[q], p.nwave[q], ts, fr, fi);
            if (before) wave_value(p.wave[q], p.nwave[q], ts - 1, pr, pi);
            const T vn = (fr * ar) - (fi * ai);
            const T vo = (pr * ar) - (pi * ai);
            out = m * (vn - ((T(1) - sp) / (T(1) + sp)) * vo);
        }
    }
    p.dw[t] = out;
}

// End of a block of steps: advance the device timestep (one lane writes).
__global__ void tick_kernel(int64_t *clk, int64_t n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) clk[0] = clk[0] + n;
}

// Set the device timestep (a restored checkpoint's), in stream order.
__global__ void set_clock_kernel(int64_t *clk, int64_t ts) {
    if (blockIdx.x == 0 && threadIdx.x == 0) clk[0] = ts;
}

// ----------------------------------------------------------- materials
// After the E update, per cell of the material box (src/Kernels/Kernels.jl:
// 76-83): [FAST: the terms the D-eliminated update left out], chi3
// (src/Kernels/Dispersive.jl:137-157), then every pole's ADE update and the
// new total polarization (:25-105).
//
// FAST mode carries D on E (or on W where the own-axis sigma is non-zero).
// With fPD the reference computes E = m (D - fPD) (W likewise); advancing the
// carried value leaves out  dw = m ((1 - s_p)/(1 + s_p) fPD_prev - fPD_cur),
// added as A += dw (s_o == 0) or W += dw, A += (1 + s_o) dw -- the
// source-correction form with S replaced by -fPD.  chi3 rescales the stored E
// but not D: where E carries D (s_o == 0) the next step's carried value is
// short by (1 - s_p)/(1 + s_p) delta, delta = E_linear - E_stored.
template <typename T, bool FAST, int NP>
__global__ __launch_bounds__(256) void material_kernel(MatArgs<T> p) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncell = (int64_t)p.n[0] * p.n[1] * p.n[2];
    if (c >= ncell) return;
    if (p.active && !p.active[c]) return;
    const int ux = (int)(c % p.n[0]);
    const int64_t r = c / p.n[0];
    const int uy = (int)(r % p.n[1]), uz = (int)(r / p.n[1]);
    const int g[3] = {p.lo[0] + ux, p.lo[1] + uy, p.lo[2] + uz};
    const size_t idx = (size_t)g[0] + (size_t)g[1] * p.sy + (size_t)g[2] * p.sz;
    const int bcur = FAST ? (int)((*p.clk + p.koff) & 1) : 0;
    const int bnext = FAST ? 1 - bcur : 0;
    T e[NP][3], so[3];
    for (int q = 0; q < NP; ++q)
        for (int a = 0; a < 3; ++a) e[q][a] = p.E[q][a][idx];
    bool dirty = false;
    if constexpr (FAST) {
        for (int a = 0; a < 3; ++a) {
            const int prv = (a + 2) % 3;
            so[a] = p.sig[a][g[a]];
            const T sp = p.sig[prv][g[prv]];
            const T ratio = (T(1) - sp) / (T(1) + sp);
            const T m = p.minv[a] ? p.minv[a][idx] : p.minv_s;
            for (int q = 0; q < NP; ++q) {
                if (p.delta[0][0] && so[a] == T(0)) e[q][a] = e[q][a] + ratio * p.delta[q][a][c];
                if (p.npole) {
                    const T dw = m * (ratio * p.fpd[bnext][q][a][c] - p.fpd[bcur][q][a][c]);
                    if (so[a] == T(0)) {
                        e[q][a] = e[q][a] + dw;
                    } else {
                        p.W[q][a][idx] = p.W[q][a][idx] + dw;
                        e[q][a] = e[q][a] + (T(1) + so[a]) * dw;
                    }
                }
            }
        }
        dirty = p.delta[0][0] || p.npole;
        if (dirty) {
            // the wrap ran before this pass: refresh the ghost copies of a
            // corrected boundary cell (linear field, as the reference wraps
            // before chi3 and P)
            for (int ax = 0; ax < 3; ++ax) {
                if (!p.per[ax]) continue;
                for (int side = 0; side < 2; ++side) {
                    if (g[ax] != (side == 0 ? p.N[ax] : 1)) continue;
                    int gd[3] = {g[0], g[1], g[2]};
                    gd[ax] = side == 0 ? 0 : p.N[ax] + 1;
                    const size_t di = (size_t)gd[0] + (size_t)gd[1] * p.sy + (size_t)gd[2] * p.sz;
                    const double *ph = side == 0 ? p.rev[ax] : p.fwd[ax];
                    for (int a = 0; a < 3; ++a) {
                        if constexpr (NP == 2) {
                            const double re = (double)e[0][a], im = (double)e[1][a];
                            p.E[0][a][di] = (T)(re * ph[0] - im * ph[1]);
                            p.E[1][a][di] = (T)(re * ph[1] + im * ph[0]);
                        } else {
                            p.E[0][a][di] = e[0][a];
                        }
                    }
                }
            }
        }
    }
    if (p.chi3) {
        const T c3 = p.chi3[c];
        if (c3 != T(0)) {
            // real(e conj(e)) per component = re*re + im*im, summed x, y, z
            T m2[3];
            for (int a = 0; a < 3; ++a) {
                m2[a] = e[0][a] * e[0][a];
                if constexpr (NP == 2) m2[a] = m2[a] + e[1][a] * e[1][a];
            }
            const T esq = m2[0] + m2[1] + m2[2];
            const T corr = T(1) / (T(1) + c3 * esq);
            for (int q = 0; q < NP; ++q)
                for (int a = 0; a < 3; ++a) {
                    const T en = e[q][a] * corr;
                    if constexpr (FAST) {
                        if (p.delta[0][0]) p.delta[q][a][c] = so[a] == T(0) ? e[q][a] - en : T(0);
                    }
                    e[q][a] = en;
                }
            dirty = true;
        }
    }
    if (dirty)
        for (int q = 0; q < NP; ++q)
            for (int a = 0; a < 3; ++a) p.E[q][a][idx] = e[q][a];
    if (p.npole) {
        T sum[NP][3];
        for (int q = 0; q < NP; ++q)
            for (int a = 0; a < 3; ++a) sum[q][a] = T(0);
        for (int k = 0; k < p.npole; ++k) {
            const PoleDev<T> &pl = p.poles[k];
            const int kx = g[0] - pl.lo[0], ky = g[1] - pl.lo[1], kz = g[2] - pl.lo[2];
            if (kx < 0 || ky < 0 || kz < 0 || kx >= pl.n[0] || ky >= pl.n[1] || kz >= pl.n[2]) {
                // outside the pole's box its P is zero: fPD += 0
                for (int q = 0; q < NP; ++q)
                    for (int a = 0; a < 3; ++a) sum[q][a] = sum[q][a] + T(0);
                continue;
            }
            const size_t kc = (size_t)kx + (size_t)pl.n[0] * ((size_t)ky + (size_t)pl.n[1] * (size_t)kz);
            const T s = pl.sigma[kc];
            for (int q = 0; q < NP; ++q)
                for (int a = 0; a < 3; ++a) {
                    T pn = pl.P[q][a][kc];
                    if (s != T(0)) {
                        const T pc = pn;
                        pn = pl.g1i * ((pl.coefp * pc - pl.g1 * pl.Pp[q][a][kc]) + (pl.drive * s) * e[q][a]);
                        pl.P[q][a][kc] = pn;
                        pl.Pp[q][a][kc] = pc;
                    }
                    sum[q][a] = sum[q][a] + pn;
                }
        }
        for (int q = 0; q < NP; ++q)
            for (int a = 0; a < 3; ++a) p.fpd[bnext][q][a][c] = sum[q][a];
    }
}

// The material box's active cells (see MatArgs::active), once at set-up.
template <typename T>
__global__ __launch_bounds__(256) void material_mask_kernel(MatArgs<T> p, uint8_t *active) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncell = (int64_t)p.n[0] * p.n[1] * p.n[2];
    if (c >= ncell) return;
    const int ux = (int)(c % p.n[0]);
    const int64_t r = c / p.n[0];
    const int uy = (int)(r % p.n[1]), uz = (int)(r / p.n[1]);
    const int g[3] = {p.lo[0] + ux, p.lo[1] + uy, p.lo[2] + uz};
    bool on = p.chi3 && p.chi3[c] != T(0);
    for (int k = 0; k < p.npole && !on; ++k) {
        const PoleDev<T> &pl = p.poles[k];
        const int kx = g[0] - pl.lo[0], ky = g[1] - pl.lo[1], kz = g[2] - pl.lo[2];
        if (kx < 0 || ky < 0 || kz < 0 || kx >= pl.n[0] || ky >= pl.n[1] || kz >= pl.n[2]) continue;
        on = pl.sigma[(size_t)kx + (size_t)pl.n[0] * ((size_t)ky + (size_t)pl.n[1] * (size_t)kz)] != T(0);
    }
    active[c] = on ? 1 : 0;
}

// Periodic / Bloch wrap of one field group along one axis
// (src/Chunking.jl:1755-1768, :2138-2172): ghost 0 <- plane N * rev,
// ghost N+1 <- plane 1 * fwd, transverse grid indices 1..N.  The phase is
// applied in complex float64 and rounded to T; a unit phase copies.
template <typename T>
struct WrapArgs {
    T *F[2][3];
    int nparts;
    int ax;
    int N[3];
    int64_t sy, sz;
    double fwd[2], rev[2];
    int unit;
};

template <typename T>
__global__ __launch_bounds__(256) void wrap_kernel(WrapArgs<T> w) {
    // the face's two axes, lower one fastest (x whenever the face holds it,
    // so that consecutive lanes touch consecutive elements of a row)
    const int ax = w.ax;
    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    const int n1 = w.N[t1], n2 = w.N[t2];
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);  // faces x 6 < 2^31 (host-checked)
    if (c >= n1 * n2 * 6) return;
    const int i = c % n1;
    int r = c / n1;
    const int j = r % n2;
    r /= n2;
    const int comp = r % 3, side = r / 3;
    int gs[3], gd[3];
    gs[ax] = side == 0 ? w.N[ax] : 1;
    gd[ax] = side == 0 ? 0 : w.N[ax] + 1;
    gs[t1] = gd[t1] = i + 1;
    gs[t2] = gd[t2] = j + 1;
    const size_t si = (size_t)gs[0] + (size_t)gs[1] * w.sy + (size_t)gs[2] * w.sz;
    const size_t di = (size_t)gd[0] + (size_t)gd[1] * w.sy + (size_t)gd[2] * w.sz;
    if (w.nparts == 1 || w.unit) {
        for (int q = 0; q < w.nparts; ++q) w.F[q][comp][di] = w.F[q][comp][si];
    } else {
        const double *ph = side == 0 ? w.rev : w.fwd;
        const double re = (double)w.F[0][comp][si], im = (double)w.F[1][comp][si];
        w.F[0][comp][di] = (T)(re * ph[0] - im * ph[1]);
        w.F[1][comp][di] = (T)(re * ph[1] + im * ph[0]);
    }
}

// ----------------------------------------------------------- monitors
// DFT accumulation of one monitor (src/Monitors/Monitors.jl:361-379):
// acc[cell, k] += (dt * exp(i * f_k * theta)) * field[cell].  The phase
// weights are formed once per block in float64 and kept in LDS; each
// accumulator update is formed in float64 and stored in ACC.
template <typename T>
struct DftMon {
    void *acc;              // complex (re, im) per cell and frequency, ACC
    const T *f, *fim;       // field (imaginary part with complex fields, else null)
    const double *freqs;
    int s[3], d[3];
    int nf, dec;
    int first;              // first block
    // FAST-mode D/B: value = (W where the own-axis sigma is non-zero, else
    // the field) / inverse material; rebuild = 0 reads f as it is
    int rebuild;
    const T *w, *wim;
    const T *minv;          // per-voxel (padded) or null: minv_s
    T minv_s;
    const T *sig_own;       // sampled sigma along the own axis (N+2)
    int own;
};

template <typename T>
struct DftArgs {
    int nmon;
    DftMon<T> mon[MAX_DFT];
    const int64_t *clk;
    int koff;
    T dt;
    int electric;           // E monitors sample at t + dt/2
    int n0, n1, n2;
    int64_t sy, sz;
    // x ghost planes not filled (periodic x folded into the updates): a
    // sample on x = 0 or x = N + 1 is formed from x = N or x = 1 as the wrap
    // would have formed it
    int perx, xunit;
    double xfwd[2], xrev[2];
};

// 2 pi t of a DFT sample at step ts: E monitors sample at t + dt/2, H at t;
// time_fac = complex_backend_number(im * 2 * pi * time) (src/Monitors/Monitors.jl:287)
template <typename T>
__device__ __forceinline__ double dft_theta(int64_t ts, T dt, int electric) {
    double t = step_time(ts, dt);
    if (electric) t = t + (double)(dt / T(2));
    return (double)(T)(6.283185307179586 * t);
}

// All DFT monitors of one half step in one launch; a monitor not due this
// step (ts % decimation != 0) exits (src/Kernels/Kernels.jl:465-497).
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void dft_kernel(DftArgs<T> p) {
    extern __shared__ double wsh[];
    int q = 0;
    const int b = blockIdx.x;
    while (q + 1 < p.nmon && b >= p.mon[q + 1].first) ++q;
    const DftMon<T> &m = p.mon[q];
    const int64_t ts = *p.clk + p.koff;
    if (m.dec > 1 && (ts % m.dec) != 0) return;
    const double theta = dft_theta(ts, p.dt, p.electric);
    const double dts = (double)p.dt;
    for (int k = threadIdx.x; k < m.nf; k += blockDim.x) {
        double sn, cs;
        sincos(m.freqs[k] * theta, &sn, &cs);
        wsh[2 * k] = dts * cs;
        wsh[2 * k + 1] = dts * sn;
    }
    __syncthreads();
    const int d0 = m.d[0], d1 = m.d[1], d2 = m.d[2];
    const int64_t ncell = (int64_t)d0 * d1 * d2;
    const int64_t cell = (int64_t)(b - m.first) * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    int ix = (int)(cell % d0);
    int64_t r = cell / d0;
    int iy = (int)(r % d1);
    int iz = (int)(r / d1);
    const int gx = m.s[0] + ix, gy = m.s[1] + iy, gz = m.s[2] + iz;
    // past grid index N+1 the reference's array holds an always-zero ghost
    const bool in = gx <= p.n0 + 1 && gy <= p.n1 + 1 && gz <= p.n2 + 1;
    const size_t fi = (size_t)gx + (size_t)gy * p.sy + (size_t)gz * p.sz;
    const T *src = m.f, *srci = m.fim;
    T mi = T(1);
    if (m.rebuild && in) {
        const int go = m.own == 0 ? gx : (m.own == 1 ? gy : gz);
        if (m.sig_own[go] != T(0)) {
            src = m.w;
            srci = m.wim;
        }
        mi = m.minv ? m.minv[fi] : m.minv_s;
    }
    double v = in ? (double)(m.rebuild ? src[fi] / mi : src[fi]) : 0.0;
    double u = in && m.fim ? (double)(m.rebuild ? srci[fi] / mi : srci[fi]) : 0.0;
    if (p.perx && !m.rebuild && (gx == 0 || gx == p.n0 + 1) && gy >= 1 && gy <= p.n1 && gz >= 1 && gz <= p.n2) {
        const bool hi = gx == p.n0 + 1;
        const size_t si = hi ? fi - (size_t)p.n0 : fi + (size_t)p.n0;
        const T fr = src[si], fm = m.fim ? srci[si] : T(0);
        if (m.fim && !p.xunit) {
            const double *ph = hi ? p.xfwd : p.xrev;
            const double re = (double)fr, im = (double)fm;
            v = (double)(T)(re * ph[0] - im * ph[1]);
            u = (double)(T)(re * ph[1] + im * ph[0]);
        } else {
            v = (double)fr;
            u = (double)fm;
        }
    }
    ACC *a = (ACC *)m.acc + 2 * cell;
    if (m.fim) {
        // complex field: (w_r + i w_i)(v + i u)
        for (int k = 0; k < m.nf; ++k, a += 2 * ncell) {
            const double wr = wsh[2 * k], wi = wsh[2 * k + 1];
            double re = (double)a[0] + (wr * v - wi * u);
            double im = (double)a[1] + (wr * u + wi * v);
            a[0] = (ACC)re;
            a[1] = (ACC)im;
        }
        return;
    }
    for (int k = 0; k < m.nf; ++k, a += 2 * ncell) {
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

// Mode overlaps of a planar monitor (src/Monitors/ModeMonitor.jl:345-507):
// one block per frequency; after the normal-axis average of each DFT (as in
// flux_kernel), with (e1, e2, h1, h2) the simulation's and (m1, m2, n1, n2)
// the mode's tangential E and H at the same samples,
//   s = e1 conj(n2) - e2 conj(n1)      (E_sim x H_mode*) . n
//   t = conj(m1) h2 - conj(m2) h1      (E_mode* x H_sim) . n
// out[4k..4k+3] = sum (s + t) dA, sum (s - t) dA as (re, im) pairs.  The
// mode is mode[((q * nf + k) * n2 + i2) * n1 + i1] (re, im), q = m1, m2, n1, n2.
template <typename ACC>
__global__ __launch_bounds__(256) void mode_overlap_kernel(FluxMon e1, FluxMon e2, FluxMon h1, FluxMon h2,
                                                           int ax, int t0, int t1, int n1, int n2, int nf,
                                                           const double *mode, double dA, double *out) {
    const int k = blockIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const FluxMon *m[4] = {&e1, &e2, &h1, &h2};
    const int64_t plane = (int64_t)n1 * n2;
    for (int c = threadIdx.x; c < n1 * n2; c += blockDim.x) {
        const int i1 = c % n1, i2 = c / n1;
        double re[4], im[4];
        for (int q = 0; q < 4; ++q) {
            const FluxMon &M = *m[q];
            const ACC *A = (const ACC *)M.acc;
            int ii[3];
            ii[ax] = 0;
            ii[t0] = i1;
            ii[t1] = i2;
            const int64_t nc = (int64_t)M.d[0] * M.d[1] * M.d[2];
            const int64_t lin = ii[0] + (int64_t)M.d[0] * (ii[1] + (int64_t)M.d[1] * ii[2]);
            const ACC *p0 = A + 2 * (lin + nc * k);
            double r = (double)p0[0], i = (double)p0[1];
            if (M.d[ax] >= 2) {
                const int64_t step = ax == 0 ? 1 : (ax == 1 ? (int64_t)M.d[0] : (int64_t)M.d[0] * M.d[1]);
                const ACC *p1 = p0 + 2 * step;
                r = (r + (double)p1[0]) / 2;
                i = (i + (double)p1[1]) / 2;
            }
            re[q] = r;
            im[q] = i;
        }
        double mr[4], mi[4];
        for (int q = 0; q < 4; ++q) {
            const double *pm = mode + 2 * ((int64_t)(q * nf + k) * plane + c);
            mr[q] = pm[0];
            mi[q] = pm[1];
        }
        // s = e1 conj(n2) - e2 conj(n1)
        const double sr = (re[0] * mr[3] + im[0] * mi[3]) - (re[1] * mr[2] + im[1] * mi[2]);
        const double si = (im[0] * mr[3] - re[0] * mi[3]) - (im[1] * mr[2] - re[1] * mi[2]);
        // t = conj(m1) h2 - conj(m2) h1
        const double tr = (mr[0] * re[3] + mi[0] * im[3]) - (mr[1] * re[2] + mi[1] * im[2]);
        const double ti = (mr[0] * im[3] - mi[0] * re[3]) - (mr[1] * im[2] - mi[1] * re[2]);
        acc[0] += (sr + tr) * dA;
        acc[1] += (si + ti) * dA;
        acc[2] += (sr - tr) * dA;
        acc[3] += (si - ti) * dA;
    }
    for (int q = 0; q < 4; ++q) {
        const double v = block_sum(acc[q]);
        if (threadIdx.x == 0) out[4 * k + q] = v;
        __syncthreads();   // block_sum's LDS partials are reused by the next sum
    }
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

// ------------------------------------------------------------ near-to-far
// Surface-equivalence projection of the tangential DFT fields on a plane to
// observation points (src/Monitors/Near2Far.jl:103-246 kernels, :380-550
// driver).  All arithmetic is FP64.

struct N2FGeom {
    double base[4][3];       // position of sample (0, 0) of e1, e2, h1, h2
    double d1, d2, dA, ns;   // tangential spacings (lower axis first), cell area, normal sign
    double eps, mu, Z;       // medium; Z = sqrt(mu / eps)
};

constexpr double N2F_TWO_PI = 6.283185307179586;    // Float64(2 pi)
constexpr double N2F_FOUR_PI = 12.566370614359172;  // Float64(4 pi)
constexpr double N2F_HALF_PI = 1.5707963267948966;  // Float64(pi) / 2

// green3d! (src/Monitors/Near2Far.jl:40-94): adds to EH = (Ex, Ey, Ez, Hx, Hy,
// Hz) as (re, im) pairs the field at o of a point current of component C
// (0-2: Jx..Jz, 3-5: Mx..Mz) with amplitude (fr, fi) at s.  The operations
// are the reference's, in its order (the library is built without FMA
// contraction):
//   1/(ikr) = (0, -1/kr), 1/(ikr)^2 = 1/(-(kr)^2)
//   term1 = 1 - 1/(ikr) + 1/(ikr)^2, term2 = (-1 + 3/(ikr) - 3/(ikr)^2) (p.r),
//   term3 = 1 - 1/(ikr), expfac = f (k n / (4 pi r)) exp(i (kr + pi/2)).
// The unit direction p and r x p have zero entries whose products add exact
// zeros; those terms are skipped.  UNIT (eps = mu = 1) skips the divisions by
// eps, mu and Z, which are then exact identities.
template <int C, bool UNIT>
__device__ __forceinline__ void green3d(double (&EH)[12], double xo, double yo, double zo, double sx, double sy,
                                        double sz, double k, double kn, const N2FGeom &g, double fr, double fi) {
    const double rx = xo - sx, ry = yo - sy, rz = zo - sz;
    const double r = sqrt((rx * rx + ry * ry) + rz * rz);
    if (!(r >= 1e-20)) return;
    const double h[3] = {rx / r, ry / r, rz / r};
    const double kr = k * r;
    double sn, cs;
    sincos(kr + N2F_HALF_PI, &sn, &cs);
    const double s = kn / (N2F_FOUR_PI * r);
    const double ar = fr * s, ai = fi * s;
    const double er = ar * cs - ai * sn, ei = ar * sn + ai * cs;
    const double a1 = 1.0 / kr;
    const double kr2 = -(kr * kr);
    const double t1r = 1.0 + 1.0 / kr2, t1i = a1;
    constexpr int c = C % 3;
    const double pd = h[c];
    const double t2r = (-1.0 - 3.0 / kr2) * pd, t2i = (3.0 * -a1) * pd;
    double efr = er, efi = ei;
    if (!UNIT) {
        const double m = C < 3 ? g.eps : g.mu;
        efr = er / m;
        efi = ei / m;
    }
    constexpr int D = C < 3 ? 0 : 6;  // slots of the parallel part: E for J, H for M
    constexpr int X = C < 3 ? 6 : 0;  // slots of the r x p part
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double wr = t2r * h[j], wi = t2i * h[j];
        if (j == c) {
            wr = t1r + wr;
            wi = t1i + wi;
        }
        EH[D + 2 * j] += efr * wr - efi * wi;
        EH[D + 2 * j + 1] += efr * wi + efi * wr;
    }
    // ef * term3 with term3 = (1, a1)
    const double ur = efr - efi * a1, ui = efr * a1 + efi;
    // (r x p)_{c+1} = h_{c+2}, (r x p)_{c+2} = -h_{c+1}
    constexpr int j1 = (c + 1) % 3, j2 = (c + 2) % 3;
    const double q1 = h[j2], q2 = -h[j1];
    if (C < 3) {  // + ef term3 (r x p) / Z
        double a = ur * q1, b = ui * q1, e = ur * q2, f = ui * q2;
        if (!UNIT) {
            a = a / g.Z;
            b = b / g.Z;
            e = e / g.Z;
            f = f / g.Z;
        }
        EH[X + 2 * j1] += a;
        EH[X + 2 * j1 + 1] += b;
        EH[X + 2 * j2] += e;
        EH[X + 2 * j2 + 1] += f;
    } else {      // - ef term3 (r x p) Z
        double a = -ur * q1, b = -ui * q1, e = -ur * q2, f = -ui * q2;
        if (!UNIT) {
            a = a * g.Z;
            b = b * g.Z;
            e = e * g.Z;
            f = f * g.Z;
        }
        EH[X + 2 * j1] += a;
        EH[X + 2 * j1 + 1] += b;
        EH[X + 2 * j2] += e;
        EH[X + 2 * j2 + 1] += f;
    }
}

// sin and cos of x >= 0 for the FAST projection: x = n pi/2 + t by a
// two-constant Cody-Waite reduction done with FMAs (each step rounds once, so
// |t| <= pi/4 carries an absolute error of about 1e-16), then the minimax
// kernels of fdlibm (Sun Microsystems, k_sin.c / k_cos.c, public domain) on
// [-pi/4, pi/4], evaluated with FMAs, and the quadrant of n.  x >= 2^31 pi/2
// (and non-finite x) go to the library sincos.
// a * b + c with c held in a scalar register pair: the VOP3 form, so a
// constant c is not copied into the destination first (as the two-address
// v_fmac the compiler prefers would need).
__device__ __forceinline__ double n2f_fma_s(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

__device__ __forceinline__ void n2f_sincos(double x, double *sn, double *cs) {
    if (!(x < 3.0e9)) {
        sincos(x, sn, cs);
        return;
    }
    const double n = __builtin_rint(x * 0.63661977236758134308);   // 2 / pi
    double t = __builtin_fma(-n, 1.5707963267948966, x);              // pi/2, leading double
    t = __builtin_fma(-n, 6.123233995736766e-17, t);                  // pi/2, next double
    const double z = t * t;
    double ps = n2f_fma_s(z, 1.58969099521155010221e-10, -2.50507602534068634195e-08);
    ps = n2f_fma_s(z, ps, 2.75573137070700676789e-06);
    ps = n2f_fma_s(z, ps, -1.98412698298579493134e-04);
    ps = n2f_fma_s(z, ps, 8.33333333332248946124e-03);
    ps = n2f_fma_s(z, ps, -1.66666666666666324348e-01);
    const double s = __builtin_fma(z * t, ps, t);
    double pc = n2f_fma_s(z, -1.13596475577881948265e-11, 2.08757232129817482790e-09);
    pc = n2f_fma_s(z, pc, -2.75573143513906633035e-07);
    pc = n2f_fma_s(z, pc, 2.48015872894767294178e-05);
    pc = n2f_fma_s(z, pc, -1.38888888888741095749e-03);
    pc = n2f_fma_s(z, pc, 4.16666666666666019037e-02);
    const double hz = 0.5 * z, w = 1.0 - hz;
    const double c = w + __builtin_fma(z * z, pc, (1.0 - w) - hz);
    const int q = (int)n;
    const double s1 = (q & 1) ? c : s, c1 = (q & 1) ? s : c;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

// sqrt(x) for x > 1e-200 (no denormal scaling needed) from the hardware
// reciprocal square root by the Newton/Goldschmidt sequence whose last two
// corrections round correctly (the compiler's own lowering of sqrt for
// double), so r matches an IEEE sqrt; the refined half-reciprocal gives 1/r,
// finished by one more Newton step (relative error ~1e-16).
__device__ __forceinline__ double n2f_sqrt_rcp(double x, double *inv) {
    const double y0 = __builtin_amdgcn_rsq(x);
    double g = x * y0, h = 0.5 * y0;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    const double q = 2.0 * h;
    *inv = __builtin_fma(q, __builtin_fma(-g, q, 1.0), q);
    return g;
}

// green3d! with the same terms, evaluated for speed: r, kr and the phase
// argument kr + pi/2 are formed exactly as in green3d (at r ~ 1e6 one ulp of r
// moves the phase by ~1e-9); the amplitudes use 1/r, powers of 1/(kr) by
// products, the inverses of eps, mu, Z, and FMAs.  The source amplitude
// (fr, fi) already carries ns dA and the current's sign.  Per-frequency
// constants: k, kn4 = k n / (4 pi), ik = 1/k; ieps, imu, iZ = 1/eps, 1/mu,
// 1/Z.
struct N2FFast {
    double ieps, imu, iZ;
};

template <int C, bool UNIT>
__device__ __forceinline__ void green3d_fast(double (&EH)[12], double xo, double yo, double zo, double sx, double sy,
                                             double sz, double k, double kn4, double ik, const N2FGeom &g,
                                             const N2FFast &f, double fr, double fi) {
    const double rx = xo - sx, ry = yo - sy, rz = zo - sz;
    const double r2 = (rx * rx + ry * ry) + rz * rz;
    if (!(r2 >= 1e-40)) return;
    double ir;
    double r;
    if (r2 > 1e-200) {
        r = n2f_sqrt_rcp(r2, &ir);
    } else {
        r = sqrt(r2);
        ir = 1.0 / r;
    }
    const double h[3] = {rx * ir, ry * ir, rz * ir};
    const double kr = k * r;
    double sn, cs;
    n2f_sincos(kr + N2F_HALF_PI, &sn, &cs);
    const double s = kn4 * ir;
    const double ar = fr * s, ai = fi * s;
    double er = __builtin_fma(ar, cs, -(ai * sn)), ei = __builtin_fma(ar, sn, ai * cs);
    if (!UNIT) {
        const double m = C < 3 ? f.ieps : f.imu;
        er *= m;
        ei *= m;
    }
    const double a1 = ir * ik;       // 1/(kr)
    const double b = a1 * a1;        // 1/(kr)^2
    constexpr int c = C % 3;
    const double pd = h[c];
    const double t2r = __builtin_fma(3.0, b, -1.0) * pd, t2i = (-3.0 * a1) * pd;
    constexpr int D = C < 3 ? 0 : 6;
    constexpr int X = C < 3 ? 6 : 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double wr, wi;
        if (j == c) {
            wr = __builtin_fma(t2r, h[j], 1.0 - b);
            wi = __builtin_fma(t2i, h[j], a1);
        } else {
            wr = t2r * h[j];
            wi = t2i * h[j];
        }
        EH[D + 2 * j] = __builtin_fma(er, wr, __builtin_fma(-ei, wi, EH[D + 2 * j]));
        EH[D + 2 * j + 1] = __builtin_fma(er, wi, __builtin_fma(ei, wr, EH[D + 2 * j + 1]));
    }
    double ur = __builtin_fma(-ei, a1, er), ui = __builtin_fma(er, a1, ei);
    if (!UNIT) {
        const double zf = C < 3 ? f.iZ : -g.Z;
        ur *= zf;
        ui *= zf;
    } else if (C >= 3) {
        ur = -ur;
        ui = -ui;
    }
    constexpr int j1 = (c + 1) % 3, j2 = (c + 2) % 3;
    EH[X + 2 * j1] = __builtin_fma(ur, h[j2], EH[X + 2 * j1]);
    EH[X + 2 * j1 + 1] = __builtin_fma(ui, h[j2], EH[X + 2 * j1 + 1]);
    EH[X + 2 * j2] = __builtin_fma(-ur, h[j1], EH[X + 2 * j2]);
    EH[X + 2 * j2 + 1] = __builtin_fma(-ui, h[j1], EH[X + 2 * j2 + 1]);
}

// The four equivalent currents of one surface sample, in the reference's
// order for each normal axis (src/Monitors/Near2Far.jl:147-150, :192-195,
// :238-241).  v[q] = (re, im) of e1, e2, h1, h2; p[q] = source positions.
template <int AX, bool UNIT, bool FAST>
__device__ __forceinline__ void n2f_sample(double (&EH)[12], double xo, double yo, double zo, const double (&v)[4][2],
                                           const double (&p)[4][3], double k, double kn, double kn4, double ik,
                                           const N2FGeom &g, const N2FFast &f) {
    const double ns = g.ns, mns = -ns, dA = g.dA;
    // SGN = +1: amplitude ns v dA, -1: -ns v dA.  FAST tiles hold v ns dA already.
#define KH_N2F_CALL(CC, Q, SGN)                                                                                   \
    if (FAST)                                                                                                     \
        green3d_fast<CC, UNIT>(EH, xo, yo, zo, p[Q][0], p[Q][1], p[Q][2], k, kn4, ik, g, f,                       \
                               (SGN) > 0 ? v[Q][0] : -v[Q][0], (SGN) > 0 ? v[Q][1] : -v[Q][1]);                   \
    else                                                                                                          \
        green3d<CC, UNIT>(EH, xo, yo, zo, p[Q][0], p[Q][1], p[Q][2], k, kn, g,                                    \
                          (((SGN) > 0 ? ns : mns) * v[Q][0]) * dA, (((SGN) > 0 ? ns : mns) * v[Q][1]) * dA)
    if (AX == 2) {         // J = (ns Hy, -ns Hx), M = (-ns Ey, ns Ex)
        KH_N2F_CALL(0, 3, 1);
        KH_N2F_CALL(1, 2, -1);
        KH_N2F_CALL(3, 1, -1);
        KH_N2F_CALL(4, 0, 1);
    } else if (AX == 0) {  // Jy = ns Hz, Jz = -ns Hy, My = -ns Ez, Mz = ns Ey
        KH_N2F_CALL(1, 3, 1);
        KH_N2F_CALL(2, 2, -1);
        KH_N2F_CALL(4, 1, -1);
        KH_N2F_CALL(5, 0, 1);
    } else {               // Jz = ns Hx, Jx = -ns Hz, Mz = -ns Ex, Mx = ns Ez
        KH_N2F_CALL(2, 2, 1);
        KH_N2F_CALL(0, 3, -1);
        KH_N2F_CALL(5, 0, -1);
        KH_N2F_CALL(3, 1, 1);
    }
#undef KH_N2F_CALL
}

// Surface samples of the projection, gathered once per call (frequencies
// k0 + blockIdx.y): for sample c = i1 + n1 i2 of the n1 x n2 plane,
//   vals[(blockIdx.y N + c) 8 + 2q ..] = the DFT value of plane q (e1, e2,
//     h1, h2) at frequency k0 + blockIdx.y,
//     averaged over the (up to) two planes across the normal in the
//     accumulator's type (src/Monitors/Near2Far.jl:431-474), times ns dA
//     for FAST;
//   pos[c 12 + 3q ..] = base_q + i1 d1 + i2 d2, the position of the current
//     made from plane q (src/Monitors/Near2Far.jl:141-144).
template <typename ACC, int AX, bool FAST>
__global__ __launch_bounds__(256) void n2f_gather_kernel(FluxMon e1, FluxMon e2, FluxMon h1, FluxMon h2, N2FGeom g,
                                                         int n1, int n2, int k0, double *vals, double *pos) {
    constexpr int T0 = AX == 0 ? 1 : 0, T1 = AX == 2 ? 1 : 2;
    const int N = n1 * n2;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = k0 + blockIdx.y;
    if (c >= N) return;
    const int i1 = c % n1, i2 = c / n1;
    const FluxMon *m[4] = {&e1, &e2, &h1, &h2};
    const double sc = FAST ? g.ns * g.dA : 1.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const FluxMon &M = *m[q];
        const ACC *A = (const ACC *)M.acc;
        const int64_t lin = (AX == 0 ? (int64_t)i1 * M.d[0] + (int64_t)i2 * M.d[0] * M.d[1]
                             : AX == 1 ? (int64_t)i1 + (int64_t)i2 * M.d[0] * M.d[1]
                                       : (int64_t)i1 + (int64_t)i2 * M.d[0]);
        const int64_t nc = (int64_t)M.d[0] * M.d[1] * M.d[2];
        const ACC *p0 = A + 2 * (lin + nc * k);
        ACC r = p0[0], im = p0[1];
        if (M.d[AX] >= 2) {
            const int64_t step = AX == 0 ? 1 : (AX == 1 ? (int64_t)M.d[0] : (int64_t)M.d[0] * M.d[1]);
            r = (r + p0[2 * step]) / ACC(2);
            im = (im + p0[2 * step + 1]) / ACC(2);
        }
        double *w = vals + 8 * ((int64_t)blockIdx.y * N + c) + 2 * q;
        w[0] = FAST ? (double)r * sc : (double)r;
        w[1] = FAST ? (double)im * sc : (double)im;
        if (blockIdx.y == 0) {
            double *x = pos + 12 * (int64_t)c + 3 * q;
            x[AX] = g.base[q][AX];
            x[T0] = g.base[q][T0] + (double)i1 * g.d1;
            x[T1] = g.base[q][T1] + (double)i2 * g.d2;
        }
    }
}

// One thread per (observation point, frequency k0 + blockIdx.y); FAST selects
// green3d_fast.  The block's threads share a frequency and walk the samples
// [chunk * blockIdx.z, + chunk) of the gathered surface through an LDS tile
// (20 doubles a sample: four values, four positions), each sample read once
// per block.  The block's partial field goes to
// out[((z nf + k) 6 + j) n_obs + o] (re, im).
template <int AX, bool UNIT, bool FAST>
__global__ __launch_bounds__(256) void near2far_kernel(const double *vals, const double *pos, N2FGeom g,
                                                       const double *obs, const double *freqs, int n_obs, int nf,
                                                       int k0, int N, int chunk, double *out) {
    __shared__ double tile[256 * 20];
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = k0 + blockIdx.y, z = blockIdx.z;
    const int c0 = z * chunk, c1 = min(N, c0 + chunk);
    double xo = 0.0, yo = 0.0, zo = 0.0;
    if (o < n_obs) {
        xo = obs[3 * o];
        yo = obs[3 * o + 1];
        zo = obs[3 * o + 2];
    }
    const double n = sqrt(g.eps * g.mu);
    const double kk = N2F_TWO_PI * freqs[k] * n;
    const double kn = kk * n;
    const double kn4 = kn * (1.0 / N2F_FOUR_PI), ik = 1.0 / kk;
    const N2FFast ff = {1.0 / g.eps, 1.0 / g.mu, 1.0 / g.Z};
    const double *vk = vals + 8 * (int64_t)blockIdx.y * N;
    double EH[12] = {};
    for (int b0 = c0; b0 < c1; b0 += blockDim.x) {
        const int cnt = min((int)blockDim.x, c1 - b0);
        // coalesced copy of cnt samples: 8 values then 12 position doubles each
        for (int t = threadIdx.x; t < cnt * 20; t += blockDim.x) {
            const int s = t / 20, e = t - 20 * s;
            tile[t] = e < 8 ? vk[8 * (int64_t)(b0 + s) + e] : pos[12 * (int64_t)(b0 + s) + (e - 8)];
        }
        __syncthreads();
        if (o < n_obs) {
            for (int s = 0; s < cnt; ++s) {
                const double *T = tile + 20 * s;
                double v[4][2], p[4][3];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    v[q][0] = T[2 * q];
                    v[q][1] = T[2 * q + 1];
                    p[q][0] = T[8 + 3 * q];
                    p[q][1] = T[9 + 3 * q];
                    p[q][2] = T[10 + 3 * q];
                }
                n2f_sample<AX, UNIT, FAST>(EH, xo, yo, zo, v, p, kk, kn, kn4, ik, g, ff);
            }
        }
        __syncthreads();
    }
    if (o < n_obs) {
        const int64_t base = ((int64_t)z * nf + k) * 6;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double *w = out + 2 * ((base + j) * n_obs + o);
            w[0] = EH[2 * j];
            w[1] = EH[2 * j + 1];
        }
    }
}

// Sum of the per-split partial fields in split order: out[i] = sum_z part[z m + i].
__global__ __launch_bounds__(256) void n2f_reduce_kernel(const double *part, int splits, int64_t m, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double r = 0.0, im = 0.0;
    for (int z = 0; z < splits; ++z) {
        r += part[2 * ((int64_t)z * m + i)];
        im += part[2 * ((int64_t)z * m + i) + 1];
    }
    out[2 * i] = r;
    out[2 * i + 1] = im;
}

// Spectral samples of the four tangential planes for the layered projection
// (src/Monitors/Near2Far.jl:805-840): the reference zero-pads each plane to
// p1 x p2 (4x), FFTs it and interpolates the sorted spectrum bilinearly.  Only
// the bins next to the observation directions are read, so they are summed
// directly: out[b][q] = sum_{i1 < n1, i2 < n2} F_q(i1, i2)
// exp(-2 pi i (m1 i1 / p1 + m2 i2 / p2)) for bin b = (m1, m2), F_q the plane
// averaged across the normal in the accumulator's type.  One block per bin.
template <typename ACC, int AX>
__global__ __launch_bounds__(256) void n2f_spectrum_kernel(FluxMon e1, FluxMon e2, FluxMon h1, FluxMon h2,
                                                           int n1, int n2, int k, int p1, int p2, const int *bins,
                                                           double *out) {
    const int b = blockIdx.x;
    const int m1 = bins[2 * b], m2 = bins[2 * b + 1];
    const FluxMon *m[4] = {&e1, &e2, &h1, &h2};
    double acc[8] = {};
    const int N = n1 * n2;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        const int i1 = c % n1, i2 = c / n1;
        // exact integer phase reduction: (m i) mod p in [0, p)
        int64_t r1 = ((int64_t)m1 * i1) % p1, r2 = ((int64_t)m2 * i2) % p2;
        if (r1 < 0) r1 += p1;
        if (r2 < 0) r2 += p2;
        double sn, cs;
        sincospi(-2.0 * ((double)r1 / p1 + (double)r2 / p2), &sn, &cs);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const FluxMon &M = *m[q];
            const ACC *A = (const ACC *)M.acc;
            const int64_t lin = (AX == 0 ? (int64_t)i1 * M.d[0] + (int64_t)i2 * M.d[0] * M.d[1]
                                 : AX == 1 ? (int64_t)i1 + (int64_t)i2 * M.d[0] * M.d[1]
                                           : (int64_t)i1 + (int64_t)i2 * M.d[0]);
            const int64_t nc = (int64_t)M.d[0] * M.d[1] * M.d[2];
            const ACC *p0 = A + 2 * (lin + nc * k);
            ACC r = p0[0], im = p0[1];
            if (M.d[AX] >= 2) {
                const int64_t step = AX == 0 ? 1 : (AX == 1 ? (int64_t)M.d[0] : (int64_t)M.d[0] * M.d[1]);
                r = (r + p0[2 * step]) / ACC(2);
                im = (im + p0[2 * step + 1]) / ACC(2);
            }
            const double vr = (double)r, vi = (double)im;
            acc[2 * q] += vr * cs - vi * sn;
            acc[2 * q + 1] += vr * sn + vi * cs;
        }
    }
    for (int q = 0; q < 8; ++q) {
        const double v = block_sum(acc[q]);
        if (threadIdx.x == 0) out[8 * b + q] = v;
        __syncthreads();   // block_sum's LDS partials are reused by the next sum
    }
}

// ------------------------------------------------------------ material raster
// Geometry rasterisation and subpixel smoothing on the device
// (src/Geometry.jl:450-663 painting, :792-955 smoothing).  Shapes: 0 box
// (full size), 1 sphere (radius p[0]), 2 cylinder (radius p[0], height p[1],
// axis).  value[c] is the relative permittivity (c < 3) or permeability of
// component c, NaN where the object does not set it.
struct RasterObj {
    int shape, axis;
    double c[3], p[3];
    double value[6];
    const double *v;  // prism: nv (x, y) vertices about c, counter-clockwise (device copy)
    int nv;
};

// Prism membership: |z - cz| <= h/2 and left of (or on) every edge.
__device__ __forceinline__ bool prism_contains(const RasterObj &o, const double *P) {
    const double x = P[0] - o.c[0], y = P[1] - o.c[1];
    if (!(fabs(P[2] - o.c[2]) <= o.p[0] / 2)) return false;
    for (int i = 0; i < o.nv; ++i) {
        const int j = i + 1 == o.nv ? 0 : i + 1;
        const double ax = o.v[2 * i], ay = o.v[2 * i + 1], bx = o.v[2 * j], by = o.v[2 * j + 1];
        if (!((bx - ax) * (y - ay) - (by - ay) * (x - ax) >= 0)) return false;
    }
    return true;
}

// The object's bounding box [lo, hi] (the host painting's index range) and
// the shape's own test, with the host's arithmetic.
__device__ __forceinline__ bool raster_holds(const RasterObj &o, double x, double y, double z) {
    const double P[3] = {x, y, z};
    double lo[3], hi[3];
    if (o.shape == 0) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = o.c[a] - o.p[a] / 2;
            hi[a] = o.c[a] + o.p[a] / 2;
        }
    } else if (o.shape == 3) {
        double mn[2] = {o.v[0], o.v[1]}, mx[2] = {o.v[0], o.v[1]};
        for (int i = 1; i < o.nv; ++i)
            for (int a = 0; a < 2; ++a) {
                mn[a] = fmin(mn[a], o.v[2 * i + a]);
                mx[a] = fmax(mx[a], o.v[2 * i + a]);
            }
        for (int a = 0; a < 2; ++a) {
            lo[a] = o.c[a] + mn[a];
            hi[a] = o.c[a] + mx[a];
        }
        lo[2] = o.c[2] + -o.p[0] / 2;
        hi[2] = o.c[2] + o.p[0] / 2;
    } else if (o.shape == 1) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = o.c[a] - o.p[0];
            hi[a] = o.c[a] + o.p[0];
        }
    } else {
        for (int a = 0; a < 3; ++a) {
            const double h = a == o.axis ? o.p[1] / 2 : o.p[0];
            lo[a] = o.c[a] - h;
            hi[a] = o.c[a] + h;
        }
    }
    for (int a = 0; a < 3; ++a)
        if (!(P[a] >= lo[a] && P[a] <= hi[a])) return false;
    if (o.shape == 0) {
        for (int a = 0; a < 3; ++a)
            if (!(fabs(P[a] - o.c[a]) <= o.p[a] / 2)) return false;
        return true;
    }
    if (o.shape == 3) return prism_contains(o, P);
    if (o.shape == 1) {
        const double dx = x - o.c[0], dy = y - o.c[1], dz = z - o.c[2];
        return (dx * dx + dy * dy) + dz * dz <= o.p[0] * o.p[0];
    }
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a)
        if (a != o.axis) {
            const double d = P[a] - o.c[a];
            r2 = r2 + d * d;
        }
    return r2 <= o.p[0] * o.p[0] && fabs(P[o.axis] - o.c[o.axis]) <= o.p[1] / 2;
}

// The shape's own test without the bounding box (the smoothing's
// `contains`).
__device__ __forceinline__ bool raster_contains(const RasterObj &o, const double *P) {
    if (o.shape == 3) return prism_contains(o, P);
    if (o.shape == 0) {
        for (int a = 0; a < 3; ++a)
            if (!(fabs(P[a] - o.c[a]) <= o.p[a] / 2)) return false;
        return true;
    }
    if (o.shape == 1) {
        const double dx = P[0] - o.c[0], dy = P[1] - o.c[1], dz = P[2] - o.c[2];
        return (dx * dx + dy * dy) + dz * dz <= o.p[0] * o.p[0];
    }
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a)
        if (a != o.axis) {
            const double d = P[a] - o.c[a];
            r2 = r2 + d * d;
        }
    return r2 <= o.p[0] * o.p[0] && fabs(P[o.axis] - o.c[o.axis]) <= o.p[1] / 2;
}

// out[i + n0 (j + n1 k)] = 1 / value of the first object (list order) that
// holds (xs[i], ys[j], zs[k]) and sets component comp, else 1 -- the
// reference paints objects last to first.
template <typename T>
__global__ __launch_bounds__(256) void raster_kernel(const RasterObj *objs, int n_objs, int comp, int n0, int n1,
                                                     int n2, const double *xs, const double *ys, const double *zs,
                                                     T *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)n0 * n1 * n2;
    if (t >= n) return;
    const int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((int64_t)n0 * n1));
    const double x = xs[i], y = ys[j], z = zs[k];
    T v = T(1);
    for (int o = 0; o < n_objs; ++o) {
        const double p = objs[o].value[comp];
        if (p != p) continue;
        if (raster_holds(objs[o], x, y, z)) {
            v = T(1) / T(p);
            break;
        }
    }
    out[t] = v;
}

// Nearest surface point and outward normal of a shape (GeometryPrimitives'
// surfpt_nearby, restated): box -- inside, the nearest face (first axis of
// the largest |L| - h), outside, the clamp onto the box with the normal along
// P - clamp; sphere -- radial; cylinder -- the side or a cap inside, the
// clamp outside.
__device__ __forceinline__ void raster_normalize(double *v) {
    const double n = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (n > 0) {
        v[0] = v[0] / n;
        v[1] = v[1] / n;
        v[2] = v[2] / n;
    } else {
        v[0] = 0.0;
        v[1] = 0.0;
        v[2] = 1.0;
    }
}

// Prism: the cylinder's rule with the polygon in place of the circle (in
// plane: largest edge distance inside, nearest boundary point outside).
__device__ __forceinline__ void prism_surfpt(const RasterObj &o, const double *P, double *sp, double *nout) {
    const double x = P[0] - o.c[0], y = P[1] - o.c[1], ell = P[2] - o.c[2];
    double best = -INFINITY, bnx = 0.0, bny = 0.0, dmin = INFINITY, dnx = 0.0, dny = 0.0;
    for (int i = 0; i < o.nv; ++i) {
        const int j = i + 1 == o.nv ? 0 : i + 1;
        const double ax = o.v[2 * i], ay = o.v[2 * i + 1];
        const double ex = o.v[2 * j] - ax, ey = o.v[2 * j + 1] - ay;
        const double ln = sqrt(ex * ex + ey * ey);
        const double nx = ey / ln, ny = -ex / ln;
        const double sd = (x - ax) * nx + (y - ay) * ny;
        if (sd > best) {
            best = sd;
            bnx = nx;
            bny = ny;
        }
        const double t = fmin(fmax(((x - ax) * ex + (y - ay) * ey) / (ln * ln), 0.0), 1.0);
        const double px = x - (ax + t * ex), py = y - (ay + t * ey);
        const double d = sqrt(px * px + py * py);
        if (d < dmin) {
            dmin = d;
            dnx = px;
            dny = py;
        }
    }
    const bool inpoly = best <= 0;
    const double d_side = inpoly ? best : dmin;
    // a point on the boundary that rounding puts outside (best > 0 with
    // dmin == 0, e.g. on a vertex) takes the best edge's normal
    const bool edge_n = inpoly || !(dmin > 0);
    const double ns[3] = {edge_n ? bnx : dnx / dmin, edge_n ? bny : dny / dmin, 0.0};
    const double sgn = ell >= 0 ? 1.0 : -1.0;
    const double d_cap = fabs(ell) - o.p[0] / 2;
    const double nc[3] = {0.0, 0.0, sgn};
    if (d_side <= 0 && d_cap <= 0) {
        const bool side = d_side >= d_cap;
        const double d = side ? d_side : d_cap;
        for (int a = 0; a < 3; ++a) {
            const double n = side ? ns[a] : nc[a];
            sp[a] = P[a] - d * n;
            nout[a] = n;
        }
        return;
    }
    const double ds = fmax(d_side, 0.0), dc = fmax(d_cap, 0.0);
    for (int a = 0; a < 3; ++a) {
        const double delta = ds * ns[a] + dc * nc[a];
        sp[a] = P[a] - delta;
        nout[a] = delta;
    }
    raster_normalize(nout);
}

__device__ __forceinline__ void raster_surfpt(const RasterObj &o, const double *P, double *sp, double *nout) {
    if (o.shape == 3) {
        prism_surfpt(o, P, sp, nout);
        return;
    }
    if (o.shape == 0) {
        double L[3], s[3], d[3], h[3];
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
            h[a] = o.p[a] / 2;
            L[a] = P[a] - o.c[a];
            s[a] = L[a] >= 0 ? 1.0 : -1.0;
            d[a] = fabs(L[a]) - h[a];
            inside = inside && d[a] <= 0;
        }
        if (inside) {
            int k = 0;
            for (int a = 1; a < 3; ++a)
                if (d[a] > d[k]) k = a;
            for (int a = 0; a < 3; ++a) {
                sp[a] = P[a];
                nout[a] = 0.0;
            }
            sp[k] = o.c[k] + s[k] * h[k];
            nout[k] = s[k];
        } else {
            for (int a = 0; a < 3; ++a) {
                const double lo = o.c[a] - h[a], hi = o.c[a] + h[a];
                sp[a] = fmin(fmax(P[a], lo), hi);
                nout[a] = P[a] - sp[a];
            }
            raster_normalize(nout);
        }
        return;
    }
    if (o.shape == 1) {
        double v[3] = {P[0] - o.c[0], P[1] - o.c[1], P[2] - o.c[2]};
        const double n = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        if (n > 0) {
            nout[0] = v[0] / n;
            nout[1] = v[1] / n;
            nout[2] = v[2] / n;
        } else {
            nout[0] = 1.0;
            nout[1] = 0.0;
            nout[2] = 0.0;
        }
        for (int a = 0; a < 3; ++a) sp[a] = o.c[a] + o.p[0] * nout[a];
        return;
    }
    const int ax = o.axis;
    double v[3] = {P[0] - o.c[0], P[1] - o.c[1], P[2] - o.c[2]};
    const double ell = v[ax];
    double rho[3] = {v[0], v[1], v[2]};
    rho[ax] = 0.0;
    const double rn = sqrt((rho[0] * rho[0] + rho[1] * rho[1]) + rho[2] * rho[2]);
    double rhat[3] = {0.0, 0.0, 0.0};
    if (rn > 0) {
        for (int a = 0; a < 3; ++a) rhat[a] = rho[a] / rn;
    } else {
        rhat[(ax + 1) % 3] = 1.0;
    }
    const double sgn = ell >= 0 ? 1.0 : -1.0;
    const double d_side = rn - o.p[0];
    const double d_cap = fabs(ell) - o.p[1] / 2;
    double ncap[3] = {0.0, 0.0, 0.0};
    ncap[ax] = sgn;
    if (d_side <= 0 && d_cap <= 0) {
        if (d_side >= d_cap) {
            for (int a = 0; a < 3; ++a) {
                sp[a] = P[a] - d_side * rhat[a];
                nout[a] = rhat[a];
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                sp[a] = P[a] - d_cap * ncap[a];
                nout[a] = ncap[a];
            }
        }
        return;
    }
    const double ds = fmax(d_side, 0.0), dc = fmax(d_cap, 0.0);
    for (int a = 0; a < 3; ++a) {
        const double delta = ds * rhat[a] + dc * ncap[a];
        sp[a] = P[a] - delta;
        nout[a] = delta;
    }
    raster_normalize(nout);
}

// Fraction of the box [lo, hi] with (x - r0) . n <= 0: with the box mapped to
// [0, L] and the axes flipped so every used n_i > 0, the volume of
// {n . x <= t} is sum over the corners v of the k used axes of
// (-1)^|v| max(0, t - n . (v L))^k / (k! prod n_i L_i); an axis whose n_i L_i
// is below 1e-9 of the largest is dropped (it contributes its length).
__device__ __forceinline__ double raster_volfrac(const double *lo, const double *hi, const double *n,
                                                 const double *r0) {
    double L[3], an[3], span[3], t = 0.0;
    for (int a = 0; a < 3; ++a) {
        L[a] = hi[a] - lo[a];
        const double base = n[a] >= 0 ? lo[a] : hi[a];
        t = a == 0 ? n[a] * (r0[a] - base) : t + n[a] * (r0[a] - base);
        an[a] = fabs(n[a]);
        span[a] = an[a] * fabs(L[a]);
    }
    const double smax = fmax(fmax(fmax(span[0], span[1]), span[2]), 1e-300);
    int ks[3], k = 0;
    for (int a = 0; a < 3; ++a)
        if (span[a] > 1e-9 * smax) ks[k++] = a;
    if (k == 0) return t >= 0 ? 1.0 : 0.0;
    double acc = 0.0;
    for (int bits = 0; bits < (1 << k); ++bits) {
        double w = t, sign = 1.0;
        for (int j = 0; j < k; ++j)
            if (bits >> j & 1) {
                w = w - an[ks[j]] * L[ks[j]];
                sign = -sign;
            }
        const double m = w > 0 ? w : 0.0;
        const double pw = k == 1 ? m : (k == 2 ? m * m : m * m * m);
        acc = acc + sign * pw;
    }
    double pn = an[ks[0]], pl = L[ks[0]];
    for (int j = 1; j < k; ++j) {
        pn = pn * an[ks[j]];
        pl = pl * L[ks[j]];
    }
    const double fact = k == 1 ? 1.0 : (k == 2 ? 2.0 : 6.0);
    const double f = acc / ((fact * pn) * pl);
    return fmin(fmax(f, 0.0), 1.0);
}

// Subpixel smoothing of electric component `axis` (0..2), reading the
// unsmoothed raster `in`: a voxel differing (relative 1e-6) from a face
// neighbour -- the first of +x, -x, +y, -y, +z, -z -- takes the effective
// value of the plane through the nearest shape surface (first object on
// ties): fill fraction f, <eps> = f e_in + (1 - f) e_out, <1/eps> likewise;
// aniso: (1 - n_k^2) <1/eps> + n_k^2 / <eps>, else 1 / <eps>.
template <typename T>
__global__ __launch_bounds__(256) void smooth_kernel(const RasterObj *objs, int n_objs, int axis, int aniso, int n0,
                                                     int n1, int n2, const double *xs, const double *ys,
                                                     const double *zs, double hx, double hy, double hz, const T *in,
                                                     T *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)n0 * n1 * n2;
    if (t >= n) return;
    const int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((int64_t)n0 * n1));
    const T c = in[t];
    const T rtol = T(1e-6);
    const int idx[3] = {i, j, k}, dims[3] = {n0, n1, n2};
    const int64_t stride[3] = {1, n0, (int64_t)n0 * n1};
    bool found = false;
    T nb = T(0);
    for (int q = 0; q < 6 && !found; ++q) {
        const int a = q >> 1, st = (q & 1) ? -1 : 1;
        const int m = idx[a] + st;
        if (m < 0 || m >= dims[a]) continue;
        const T b = in[t + st * stride[a]];
        if (fabs(b - c) > rtol * (fabs(b) > fabs(c) ? fabs(b) : fabs(c))) {
            found = true;
            nb = b;
        }
    }
    if (!found) {
        out[t] = c;
        return;
    }
    const double P[3] = {xs[i], ys[j], zs[k]};
    double best = INFINITY, bsp[3] = {P[0], P[1], P[2]}, bn[3] = {0.0, 0.0, 1.0};
    int bi = -1;
    for (int o = 0; o < n_objs; ++o) {
        double sp[3], no[3];
        raster_surfpt(objs[o], P, sp, no);
        const double e0 = sp[0] - P[0], e1 = sp[1] - P[1], e2 = sp[2] - P[2];
        const double d2 = (e0 * e0 + e1 * e1) + e2 * e2;
        if (d2 < best) {
            best = d2;
            for (int a = 0; a < 3; ++a) {
                bsp[a] = sp[a];
                bn[a] = no[a];
            }
            bi = o;
        }
    }
    if (bi < 0) {
        out[t] = c;
        return;
    }
    raster_normalize(bn);
    const double lo[3] = {P[0] - hx, P[1] - hy, P[2] - hz}, hi[3] = {P[0] + hx, P[1] + hy, P[2] + hz};
    const double f = raster_volfrac(lo, hi, bn, bsp);
    const double ec = (double)(T(1) / c), en = (double)(T(1) / nb);
    const bool inside = raster_contains(objs[bi], P);
    const double es = inside ? ec : en, eb = inside ? en : ec;
    const double e_avg = f * es + (1 - f) * eb;
    const double e_harm = f / es + (1 - f) / eb;
    double val;
    if (aniso) {
        const double n2v = bn[axis] * bn[axis];
        val = (1 - n2v) * e_harm + n2v / e_avg;
    } else {
        val = 1.0 / e_avg;
    }
    out[t] = (T)val;
}

// ------------------------------------------------- time-reversed adjoint
// Discrete transpose of one half step of the sigma-cascade update
// (src/Kernels/Helpers.jl:113-139 curl stage, :323-340 material stage, the
// curls of src/Kernels/Helpers.jl:286-298): with K = dt curl(F) and
// (o, n, p) the own/next/prev-axis sigma of a component,
//   U' = ((1-n) U + K) / (1+n)          (n != 0; else U' - U reads as K)
//   T' = ((1-p) T + U' - U) / (1+p)
//   W' = m T'                            (plus sources, which the transpose drops)
//   A' = A + (1+o) W' - (1-o) W          (o != 0; else A' = W')
// Run backwards, the adjoints (lA, lT, lU, lW) of the state after the half
// step become those of the state before it, and Kbar = dJ/dK is written for
// the other group's transposed curl.  That curl is read from the other
// group's Kbar of the previous transposed half step: the transpose of a
// forward-difference curl is minus the backward-difference curl of the same
// shape, and the other way round.  Cells 1..N of every axis, as the forward
// update; ghost entries of Kbar stay zero.
template <typename T>
struct AdjSeed {
    const double *g;
    T *lam;
    const double *freqs;
    int s[3], d[3];
    int nf, dec;
    int first;
    int woff;  // the monitor's columns in a row of the step weight table
};

// The seed of monitor m at grid cell (gx, gy, gz) from its columns of the
// step weight table (row: [due, then dt cos, dt sin per frequency], as
// dft_weights_kernel wrote them): the same operations on the same values as
// adj_seed_at.
template <typename T>
__device__ __forceinline__ bool adj_seed_tab(const AdjSeed<T> &m, int gx, int gy, int gz, const double *row,
                                             double &v) {
    const int ix = gx - m.s[0], iy = gy - m.s[1], iz = gz - m.s[2];
    if (ix < 0 || iy < 0 || iz < 0 || ix >= m.d[0] || iy >= m.d[1] || iz >= m.d[2]) return false;
    const double *w = row + m.woff;
    if (w[0] == 0.0) return false;
    const int64_t ncell = (int64_t)m.d[0] * m.d[1] * m.d[2];
    const int64_t cell = (int64_t)ix + (int64_t)m.d[0] * ((int64_t)iy + (int64_t)m.d[1] * iz);
    v = 0.0;
    const double *gp = m.g + 2 * cell;
    for (int k = 0; k < m.nf; ++k, gp += 2 * ncell) v += gp[0] * w[1 + 2 * k] + gp[1] * w[2 + 2 * k];
    return true;
}

// The seed of monitor m at grid cell (gx, gy, gz) of the step at ts, as
// adjoint_seed_kernel adds it (false: the cell is not on the monitor or the
// step is not sampled).
template <typename T>
__device__ __forceinline__ bool adj_seed_at(const AdjSeed<T> &m, int gx, int gy, int gz, int64_t ts, T dt,
                                            int electric, double &v) {
    if (m.dec > 1 && (ts % m.dec) != 0) return false;
    const int ix = gx - m.s[0], iy = gy - m.s[1], iz = gz - m.s[2];
    if (ix < 0 || iy < 0 || iz < 0 || ix >= m.d[0] || iy >= m.d[1] || iz >= m.d[2]) return false;
    const int64_t ncell = (int64_t)m.d[0] * m.d[1] * m.d[2];
    const int64_t cell = (int64_t)ix + (int64_t)m.d[0] * ((int64_t)iy + (int64_t)m.d[1] * iz);
    const double theta = dft_theta(ts, dt, electric);
    const double dts = (double)dt;
    v = 0.0;
    const double *gp = m.g + 2 * cell;
    for (int k = 0; k < m.nf; ++k, gp += 2 * ncell) {
        double sn, cs;
        sincos(m.freqs[k] * theta, &sn, &cs);
        v += gp[0] * (dts * cs) + gp[1] * (dts * sn);
    }
    return true;
}

template <typename T>
struct AdjArgs {
    T *lA[3], *lT[3], *lU[3], *lW[3];
    T *Kout[3];
    const T *Kin[3];
    const T *sig[3];
    const T *minv[3];
    T minv_s, dt;
    T inv[3];
    int n[3];
    int64_t sy, sz;
    // E half step: dJ/d(eps^-1) over a box per component, from m times the
    // forward net drive of the same step (dense over the box, x fastest)
    int glo[3][3], gn[3][3];
    const T *fwd[3];
    double *grad[3];
    // the objective's seeds of this step (adjoint_seed_kernel's, added to
    // lambda_A of their component in list order before the transpose reads
    // it), ts = device clock + koff
    int nseed;
    AdjSeed<T> seed[MAX_DFT];
    const int64_t *clk;
    int64_t koff;
    // this step's row of the step weight table (null: the seeds form their
    // weights from the device clock themselves)
    const double *sw;
};

// Step weight table of DFT monitors (the E update's small monitors, the
// adjoint's seeds) for len steps at device timestep + koff0 + i (row i,
// stride doubles): per monitor, 1 where its decimation samples the step
// (else 0), then dt cos and dt sin of f_k theta for each frequency, as
// dft_kernel forms its weights.
template <typename T>
struct DftW {
    int nmon;
    const double *freqs[MAX_DFT];
    int nf[MAX_DFT], dec[MAX_DFT], woff[MAX_DFT], elec[MAX_DFT];
    const int64_t *clk;
    int64_t koff0;
    int len, stride;
    T dt;
    double *out;
};

template <typename T>
__global__ __launch_bounds__(256) void dft_weights_kernel(DftW<T> p) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)p.len * p.stride) return;
    const int i = (int)(t / p.stride), col = (int)(t % p.stride);
    int q = 0;
    while (q + 1 < p.nmon && col >= p.woff[q + 1]) ++q;
    const int j = col - p.woff[q];
    const int64_t ts = *p.clk + p.koff0 + i;
    if (j == 0) {
        p.out[t] = (p.dec[q] > 1 && (ts % p.dec[q]) != 0) ? 0.0 : 1.0;
        return;
    }
    const double theta = dft_theta(ts, p.dt, p.elec[q]);
    double sn, cs;
    sincos(p.freqs[q][(j - 1) >> 1] * theta, &sn, &cs);
    p.out[t] = (double)p.dt * ((j & 1) ? cs : sn);
}

template <typename T, bool ELEC>
__global__ __launch_bounds__(256) void adjoint_half_kernel(AdjArgs<T> p) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncell = (int64_t)p.n[0] * p.n[1] * p.n[2];
    if (c >= ncell) return;
    const int gx = 1 + (int)(c % p.n[0]);
    const int64_t r = c / p.n[0];
    const int gy = 1 + (int)(r % p.n[1]);
    const int gz = 1 + (int)(r / p.n[1]);
    const int64_t sy = p.sy, sz = p.sz;
    const size_t idx = (size_t)gx + (size_t)gy * sy + (size_t)gz * sz;
    const T *Kx = p.Kin[0], *Ky = p.Kin[1], *Kz = p.Kin[2];
    const T kx = Kx[idx], ky = Ky[idx], kz = Kz[idx];
    T add[3];
    if constexpr (ELEC) {
        // minus the E update's backward-difference curl of the H group's Kbar
        add[0] = -(p.dt * (p.inv[2] * (Ky[idx - sz] - ky) - p.inv[1] * (Kz[idx - sy] - kz)));
        add[1] = -(p.dt * (p.inv[0] * (Kz[idx - 1] - kz) - p.inv[2] * (Kx[idx - sz] - kx)));
        add[2] = -(p.dt * (p.inv[1] * (Kx[idx - sy] - kx) - p.inv[0] * (Ky[idx - 1] - ky)));
    } else {
        // minus the H update's forward-difference curl of the E group's Kbar
        add[0] = -(p.dt * (p.inv[2] * (Ky[idx + sz] - ky) - p.inv[1] * (Kz[idx + sy] - kz)));
        add[1] = -(p.dt * (p.inv[0] * (Kz[idx + 1] - kz) - p.inv[2] * (Kx[idx + sz] - kx)));
        add[2] = -(p.dt * (p.inv[1] * (Kx[idx + sy] - kx) - p.inv[0] * (Ky[idx + 1] - ky)));
    }
    const T sg[3] = {p.sig[0][gx], p.sig[1][gy], p.sig[2][gz]};
    const int gc[3] = {gx, gy, gz};
    // Every load of the lane is issued before its first store (the memory
    // counter is in order: a load issued after a store waits for the store,
    // so per-component load/store rounds would cost a round trip each).  The
    // arrays of different components and stages never overlap.
    T m[3], la[3], lw[3], lt[3], lu[3];
    int gbi[3];
    double gr[3];
    T fw[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T so = sg[a], sn = sg[(a + 1) % 3];
        m[a] = p.minv[a] ? p.minv[a][idx] : p.minv_s;
        la[a] = p.lA[a][idx];
        lw[a] = so != T(0) ? p.lW[a][idx] : T(0);
        lt[a] = p.lT[a][idx];
        lu[a] = sn != T(0) ? p.lU[a][idx] : T(0);
        gbi[a] = -1;
        gr[a] = 0.0;
        fw[a] = T(0);
        if constexpr (ELEC) {
            if (p.grad[a]) {
                int bi = 0, str = 1;
                bool in = true;
                for (int k = 0; k < 3; ++k) {
                    const int o = gc[k] - p.glo[a][k];
                    in = in && o >= 0 && o < p.gn[a][k];
                    bi += o * str;
                    str *= p.gn[a][k];
                }
                if (in) {
                    gbi[a] = bi;
                    gr[a] = p.grad[a][bi];
                    fw[a] = p.fwd[a][bi];
                }
            }
        }
    }
    if (p.nseed > 0 && p.sw) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < p.nseed; ++q) {
                double v;
                if (p.seed[q].lam == p.lA[a] && adj_seed_tab(p.seed[q], gx, gy, gz, p.sw, v))
                    la[a] = (T)((double)la[a] + v);
            }
    } else if (p.nseed > 0) {
        const int64_t ts = *p.clk + p.koff;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < p.nseed; ++q) {
                double v;
                if (p.seed[q].lam == p.lA[a] && adj_seed_at(p.seed[q], gx, gy, gz, ts, p.dt, ELEC ? 1 : 0, v))
                    la[a] = (T)((double)la[a] + v);
            }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T so = sg[a], sn = sg[(a + 1) % 3], sp = sg[(a + 2) % 3];
        const T ap = la[a] + add[a];
        T wh;
        if (so == T(0)) {
            wh = ap;
            p.lA[a][idx] = T(0);
        } else {
            wh = lw[a] + (T(1) + so) * ap;
            p.lW[a][idx] = -((T(1) - so) * ap);
            p.lA[a][idx] = ap;
        }
        if constexpr (ELEC) {
            // m enters as W' = m net: dJ/dm += What * net, with fwd = m net
            // (A' where o = 0, W' where o != 0)
            if (gbi[a] >= 0) p.grad[a][gbi[a]] = gr[a] + (double)wh * ((double)fw[a] / (double)m[a]);
        }
        const T th = lt[a] + m[a] * wh;
        p.lT[a][idx] = ((T(1) - sp) / (T(1) + sp)) * th;
        const T dl = th / (T(1) + sp);
        T kb;
        if (sn == T(0)) {
            kb = dl;
        } else {
            const T uh = lu[a] + dl;
            p.lU[a][idx] = ((T(1) - sn) / (T(1) + sn)) * uh - dl;
            kb = uh / (T(1) + sn);
        }
        p.Kout[a][idx] = kb;
    }
}

// Transpose of the DFT update at step ts: lambda += Re(conj(G_k) w_k) summed
// over frequencies, with w_k = dt exp(i f_k theta) the weight the DFT kernel
// applied and G = dJ/dRe(F) + i dJ/dIm(F) per accumulator entry (same
// layout as the accumulator).  Cells 1..N only (ghost samples are constant).
// ts = device timestep + koff (the reverse walk leaves the clock at the end
// of the interval it walks, so one launch sequence serves every interval);
// each monitor skips the steps its decimation does not sample.

template <typename T>
struct AdjSeedArgs {
    int nmon;
    AdjSeed<T> mon[MAX_DFT];
    const int64_t *clk;
    int64_t koff;
    T dt;
    int electric;
    int n0, n1, n2;
    int64_t sy, sz;
};

template <typename T>
__global__ __launch_bounds__(256) void adjoint_seed_kernel(AdjSeedArgs<T> p) {
    extern __shared__ double wsh[];
    int q = 0;
    const int b = blockIdx.x;
    while (q + 1 < p.nmon && b >= p.mon[q + 1].first) ++q;
    const AdjSeed<T> &m = p.mon[q];
    const int64_t ts = *p.clk + p.koff;
    if (m.dec > 1 && (ts % m.dec) != 0) return;
    const double theta = dft_theta(ts, p.dt, p.electric);
    const double dts = (double)p.dt;
    for (int k = threadIdx.x; k < m.nf; k += blockDim.x) {
        double sn, cs;
        sincos(m.freqs[k] * theta, &sn, &cs);
        wsh[2 * k] = dts * cs;
        wsh[2 * k + 1] = dts * sn;
    }
    __syncthreads();
    const int64_t ncell = (int64_t)m.d[0] * m.d[1] * m.d[2];
    const int64_t cell = (int64_t)(b - m.first) * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const int ix = (int)(cell % m.d[0]);
    const int64_t r = cell / m.d[0];
    const int iy = (int)(r % m.d[1]);
    const int iz = (int)(r / m.d[1]);
    const int gx = m.s[0] + ix, gy = m.s[1] + iy, gz = m.s[2] + iz;
    if (gx < 1 || gy < 1 || gz < 1 || gx > p.n0 || gy > p.n1 || gz > p.n2) return;
    double v = 0.0;
    const double *gp = m.g + 2 * cell;
    for (int k = 0; k < m.nf; ++k, gp += 2 * ncell) v += gp[0] * wsh[2 * k] + gp[1] * wsh[2 * k + 1];
    const size_t fi = (size_t)gx + (size_t)gy * p.sy + (size_t)gz * p.sz;
    m.lam[fi] = (T)((double)m.lam[fi] + v);
}

// Copy of a box of up to three fields (blockIdx.y = component), dense x
// fastest; where the component's own-axis sigma is non-zero the W stage is
// copied instead (both hold m times the net drive of the step there).
template <typename T>
struct BoxCopy {
    const T *f[3];
    const T *w[3];
    const T *sig[3];
    T *out[3];
    int lo[3][3], n[3][3];
    int64_t sy, sz;
};

template <typename T>
__global__ __launch_bounds__(256) void box_copy_kernel(BoxCopy<T> p) {
    const int a = blockIdx.y;
    if (!p.out[a]) return;
    const int64_t ncell = (int64_t)p.n[a][0] * p.n[a][1] * p.n[a][2];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int ix = (int)(c % p.n[a][0]);
    const int64_t r = c / p.n[a][0];
    const int iy = (int)(r % p.n[a][1]);
    const int iz = (int)(r / p.n[a][1]);
    const int g[3] = {p.lo[a][0] + ix, p.lo[a][1] + iy, p.lo[a][2] + iz};
    const size_t fi = (size_t)g[0] + (size_t)g[1] * p.sy + (size_t)g[2] * p.sz;
    p.out[a][c] = p.sig[a][g[a]] != T(0) ? p.w[a][fi] : p.f[a][fi];
}

// ------------------------------------------------------- state copies
// Copy of up to MAX_STATE equally sized arrays in one launch (checkpoints,
// restores): blockIdx.y selects the array; lanes move 16-byte words, then
// the last few 4-byte words of an array whose size is not a multiple of 16
// (every array starts 256-byte aligned).
constexpr int MAX_STATE = 64;
struct StateCopy {
    const void *src[MAX_STATE];
    void *dst[MAX_STATE];
    int64_t w16;  // 16-byte words per array
    int tail;     // 4-byte words after them
};

__global__ __launch_bounds__(256) void state_copy_kernel(StateCopy a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (t < a.w16) {
        reinterpret_cast<uint4 *>(a.dst[k])[t] = reinterpret_cast<const uint4 *>(a.src[k])[t];
    } else if (t < a.w16 + a.tail) {
        const int64_t o = 4 * a.w16 + (t - a.w16);
        reinterpret_cast<uint32_t *>(a.dst[k])[o] = reinterpret_cast<const uint32_t *>(a.src[k])[o];
    }
}

// ------------------------------------------------------ streaming rate
// A plain one-pass stream (measurement only): NR float32 arrays read and NW
// written, two values per lane with nontemporal accesses, blocks in linear
// order: the practical HBM ceiling for a read/write mix, against which the
// updates' measured traffic is set.
struct StreamArgs {
    const float *in[6];
    float *out[3];
    int64_t n2;  // value pairs per array
};

template <int NR, int NW>
__global__ __launch_bounds__(256) void stream_kernel(StreamArgs p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n2) return;
    float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const float2 v = ldo<float, true>(p.in[r], (size_t)(2 * i));
        acc.x = acc.x + v.x;
        acc.y = acc.y + v.y;
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) sto<float, true>(p.out[w], (size_t)(2 * i), acc);
}

}  // namespace kh
