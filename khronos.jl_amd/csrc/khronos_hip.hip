// khronos_hip.hip -- host side of libkhronos_hip (the C ABI in include/khronos_hip.h).
//
// One context = one simulation resident in HBM on the current device, one
// non-blocking HIP stream.  kh_step issues, per time step, the reference's
// sub-step order (src/Kernels/Kernels.jl:64-87): H sources are folded into the
// H update kernel, then due H-component DFT launches, E update (+E sources),
// then due E-component DFT launches.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>
#include <mutex>
#include <deque>
#include <functional>
#include <map>

#include "../../include/khronos_hip.h"
#include "kh_kernels.hpp"

namespace {

thread_local std::string g_last_error;

struct KhError : std::runtime_error {
    int code;
    KhError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            throw KhError(_e == hipErrorOutOfMemory ? KH_E_NOMEM : KH_E_DEVICE,           \
                          std::string(#expr) + ": " + hipGetErrorString(_e));             \
        }                                                                                 \
    } while (0)

template <typename F>
int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return KH_OK;
    } catch (const KhError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return KH_E_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return KH_E_INVALID;
    }
}

void require(bool c, const std::string &m, int code = KH_E_INVALID) {
    if (!c) throw KhError(code, m);
}

constexpr int ALIGN = 32;
constexpr int kh_timing_slots = KH_TIMING_SLOTS;  // elements; x index 1 of every row is aligned

struct Grid {
    int n[3];
    int64_t px, py, pz;  // padded extents (pitch in x)
    int64_t front;       // elements before grid index (0,0,0)
    int64_t total;       // elements per allocation
    void init(int nx, int ny, int nz) {
        n[0] = nx; n[1] = ny; n[2] = nz;
        px = ((int64_t)nx + 2 + ALIGN - 1) / ALIGN * ALIGN;
        py = ny + 2;
        pz = nz + 2;
        front = ALIGN - 1;
        total = front + px * py * pz + ALIGN;
    }
    int64_t sy() const { return px; }
    int64_t sz() const { return px * py; }
    int64_t lin(int gx, int gy, int gz) const { return gx + px * ((int64_t)gy + py * (int64_t)gz); }
};

void comp_count(const int n[3], int c, int out[3]) {
    // src/utils.jl:26-37
    int nn[3] = {n[0], n[1], n[2] < 1 ? 1 : n[2]};
    int a = c % 3;
    if (c < 3) {
        for (int i = 0; i < 3; ++i) out[i] = nn[i] + 1;
        out[a] = nn[a];
    } else {
        for (int i = 0; i < 3; ++i) out[i] = nn[i];
        out[a] = nn[a] + 1;
    }
}

// Device memory and streams released by destroyed contexts, kept for later
// contexts of the process.  hipFree waits for the whole device, so freeing a
// finished request's arrays would stall every other context of a serving
// loop; with the pool, kh_destroy synchronises only its own stream, and a
// context of a size seen before allocates nothing.  Blocks are matched by size
// rounded up to 64 KiB; the pool holds at most KH_POOL_BYTES (default 32 GiB,
// beyond that a released block is freed) and is emptied when an allocation
// fails for want of memory.
struct DevicePool {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void *> blocks;  // (device, rounded bytes) -> block
    std::multimap<int, hipStream_t> streams;               // device -> idle stream
    size_t held = 0;
    size_t limit = 0;
    DevicePool() {
        const char *v = getenv("KH_POOL_BYTES");
        limit = v && *v ? (size_t)strtoull(v, nullptr, 10) : ((size_t)32 << 30);
    }
};

DevicePool &dpool() {
    static DevicePool *p = new DevicePool();  // outlives every context
    return *p;
}

size_t pool_round(size_t b) { return (b + 65535) / 65536 * 65536; }

void pool_trim(int dev) {
    std::vector<void *> out;
    {
        std::lock_guard<std::mutex> lock(dpool().mu);
        auto &bl = dpool().blocks;
        for (auto it = bl.begin(); it != bl.end();) {
            if (it->first.first == dev) {
                out.push_back(it->second);
                dpool().held -= it->first.second;
                it = bl.erase(it);
            } else {
                ++it;
            }
        }
    }
    for (void *q : out) (void)hipFree(q);
}

void *pool_alloc(int dev, size_t bytes) {
    const size_t rb = pool_round(bytes);
    {
        std::lock_guard<std::mutex> lock(dpool().mu);
        auto it = dpool().blocks.find({dev, rb});
        if (it != dpool().blocks.end()) {
            void *q = it->second;
            dpool().blocks.erase(it);
            dpool().held -= rb;
            return q;
        }
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, rb);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        pool_trim(dev);
        e = hipMalloc(&q, rb);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw KhError(e == hipErrorOutOfMemory ? KH_E_NOMEM : KH_E_DEVICE,
                      std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return q;
}

// the block's last use has completed
void pool_release(int dev, void *q, size_t bytes) {
    const size_t rb = pool_round(bytes);
    {
        std::lock_guard<std::mutex> lock(dpool().mu);
        if (dpool().held + rb <= dpool().limit) {
            dpool().blocks.insert({{dev, rb}, q});
            dpool().held += rb;
            return;
        }
    }
    (void)hipFree(q);
}

hipStream_t pool_stream(int dev) {
    {
        std::lock_guard<std::mutex> lock(dpool().mu);
        auto it = dpool().streams.find(dev);
        if (it != dpool().streams.end()) {
            hipStream_t s = it->second;
            dpool().streams.erase(it);
            return s;
        }
    }
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

// the stream is idle
void pool_stream_release(int dev, hipStream_t s) {
    std::lock_guard<std::mutex> lock(dpool().mu);
    dpool().streams.insert({dev, s});
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct CtxBase {
    virtual ~CtxBase() = default;
    virtual void add_source(int comp, const int32_t *s, const int32_t *d, const void *amp, int32_t *id) = 0;
    virtual void set_waveform(int id, int64_t n, const double *v) = 0;
    virtual void add_dft(int comp, const int32_t *s, const int32_t *d, int nf, const double *f, int dec, int32_t *id) = 0;
    virtual void set_chi3(const void *chi3) = 0;
    virtual void set_eps_inv(int comp, const void *a) = 0;
    virtual void add_pole(const double *c, int is_drude, const void *sigma, int32_t *id) = 0;
    virtual void read_pole(int id, int which, int comp, void *host, size_t bytes) = 0;
    virtual void write_pole(int id, int which, int comp, const void *host, size_t bytes) = 0;
    virtual void step(int64_t n) = 0;
    virtual void set_graph(int steps) = 0;
    virtual void prepare_graphs(int64_t n) = 0;
    virtual void launch_counts(int64_t *out) = 0;
    virtual int graph_block() = 0;
    virtual void sync() = 0;
    virtual bool idle() = 0;
    virtual int64_t timestep() const = 0;
    virtual bool sources_active() const = 0;
    virtual void read_field(int comp, void *host, size_t bytes) = 0;
    virtual void write_field(int comp, const void *host, size_t bytes) = 0;
    virtual void read_dft(int id, double *host, size_t bytes) = 0;
    virtual void flux(int e1, int e2, int h1, int h2, int ax, double dA, double *out) = 0;
    virtual void dft_norm(int id, double *out) = 0;
    virtual void dft_norms(int n_out, double *out) = 0;
    virtual void mode_overlap(int e1, int e2, int h1, int h2, int ax, int n1, int n2, const double *mode,
                              double dA, double *out) = 0;
    virtual void near2far(const int32_t *ids, int ax, double ns, const double *base, double d1, double d2,
                          double dA, int n1, int n2, const double *freqs, int nf, double eps, double mu,
                          const double *obs, int n_obs, double *out) = 0;
    virtual void near2far_spectrum(const int32_t *ids, int ax, int n1, int n2, int k, int p1, int p2,
                                   const int32_t *bins, int n_bins, double *out) = 0;
    virtual void reset() = 0;
    virtual void set_checkpoints(int every) = 0;
    virtual void adjoint_gradient(int nmon, const int32_t *ids, const double *cot, const int32_t *lo,
                                  const int32_t *n, double *grad) = 0;
    virtual void set_timing(int n) = 0;
    virtual void get_timing(double *out) = 0;
    virtual void *stream() = 0;
};

template <typename T>
struct Ctx final : CtxBase {
    Grid g;
    int dev = 0;  // the device the context lives on
    int mode;
    bool acc64;
    T dt, dx[3];
    double cutoff;
    hipStream_t st = nullptr;
    std::vector<DevBuf> owned;
    // device arrays (pre-offset by g.front).  Complex fields are held as two
    // real parts with identical layouts: every real-coefficient operation of
    // the time step runs part by part; the wrap, chi3, the DFT and the
    // readout combine them.
    struct Part {
        T *E[3]{}, *H[3]{}, *D[3]{}, *B[3]{}, *UD[3]{}, *UB[3]{}, *WD[3]{}, *WB[3]{};
        T *CD[3]{}, *CB[3]{};  // C stage, components with a conductivity array
    };
    Part fp[2];
    int nparts = 1;
    bool periodic[3] = {false, false, false};
    // FAST, periodic x: the PML kernels read the x-neighbour across the
    // boundary from the opposite interior plane (xfold), so the x-axis ghost
    // fill is skipped and the x ghosts are refreshed only before a field is
    // read (xghost_stale); off when anything else reads or writes those
    // ghosts (a material box or a DFT monitor on them, a field write)
    bool xfold = false, xghost_stale = false, fields_written = false;
    double ph_fwd[3][2] = {}, ph_rev[3][2] = {};
    T *epsi[3]{}, *mui[3]{};
    T *sigB[3]{}, *sigD[3]{};
    bool hassig[2][3] = {};  // [group (0: H, 1: E)][axis]: profile given (the reference's array exists)
    // material conductivity sigma_D (Ex,Ey,Ez) / sigma_B (Hx,Hy,Hz) on the
    // padded grid, 0 outside 1..N; cbox[e] = cells where some sigma of group e
    // (0: H, 1: E) is nonzero (lo > hi: none)
    T *msig[6]{};
    int cbox[2][6] = {{1, 1, 1, 0, 0, 0}, {1, 1, 1, 0, 0, 0}};
    // per-voxel eps^-1 / mu^-1 as a byte per cell into a table of <= 256
    // values (group 0: mu, 1: eps), used by the PML kernels when it fits
    uint8_t *mid[6]{};
    T *lut[2]{};
    bool use_lut[2] = {false, false};
    T eps_s, mu_s;
    int64_t ts = 0;
    bool active = true;
    bool stepped = false;

    struct Source {
        int comp;
        int s[3], d[3];
        T *amp;
        double *wave = nullptr;  // device copy of the caller's table, (re, im) per step
        std::vector<double> host;  // host copy of the same table
        int64_t nwave = 0, cap = 0;
    };
    std::vector<Source> srcs;

    struct Dft {
        int comp;
        int s[3], d[3];
        int nf, dec;
        double *freqs;
        void *acc;
        int64_t ncell;
    };
    std::vector<Dft> dfts;

    // chi3 and dispersive poles (material pass after the E update)
    std::vector<T> chi3_host;                  // (nx, ny, nz) until the first step
    std::vector<kh::PoleDev<T>> poles;
    kh::PoleDev<T> *poles_dev = nullptr;
    T *chi3_dev = nullptr;                     // dense over the material box
    T *fpd[2][2][3] = {};                      // total polarization [buffer][part][comp] (FAST: two buffers, alternating)
    T *delta[2][3] = {};                       // FAST with chi3, [part][comp]
    bool mat = false;
    int mlo[3] = {1, 1, 1}, mn[3] = {0, 0, 0};
    int64_t mcells = 0;

    double *partials = nullptr;  // reduction partials (flux, overlap, norms)
    int64_t *clk = nullptr;     // device timestep
    static constexpr int NPART = 256;

    bool timing = false;
    struct EvPair { hipEvent_t a, b; int kind; };
    std::vector<EvPair> evs;   // recorded, not yet read
    std::vector<EvPair> pool;  // created up front so timing adds no allocation

    // Launch on the context stream; with kernel timing on, the launch itself
    // stamps an event pair at the kernel's start and end (the dispatch gap in
    // front of the kernel is not counted, as a pair recorded around the
    // launch would count it).
    template <typename... Args, typename F = void (*)(Args...)>
    void launch(int kind, F kernel, dim3 grid, dim3 blk, size_t shm, Args... args) {
        if (timing) {
            EvPair ev = take_event(kind);
            hipExtLaunchKernelGGL(kernel, grid, blk, (std::uint32_t)shm, st, ev.a, ev.b, 0, args...);
            HIPCHK(hipGetLastError());
            evs.push_back(ev);
        } else {
            hipLaunchKernelGGL(kernel, grid, blk, shm, st, args...);
            HIPCHK(hipGetLastError());
        }
    }

    EvPair take_event(int kind) {
        if (pool.empty()) {
            EvPair e{};
            HIPCHK(hipEventCreate(&e.a));
            HIPCHK(hipEventCreate(&e.b));
            pool.push_back(e);
        }
        EvPair e = pool.back();
        pool.pop_back();
        e.kind = kind;
        return e;
    }
    double tsum[kh_timing_slots] = {};
    double tcnt[kh_timing_slots] = {};

    void *dalloc(size_t bytes) {
        void *p = pool_alloc(dev, bytes);
        owned.push_back({p, bytes});
        HIPCHK(hipMemsetAsync(p, 0, bytes, st));
        return p;
    }
    T *field_alloc() {
        T *p = (T *)dalloc((size_t)g.total * sizeof(T));
        return p + g.front;
    }

    Ctx(const kh_sim_desc &d) {
        g.init(d.nx, d.ny, d.nz);
        mode = d.mode;
        acc64 = d.dft_accum_f64 != 0;
        dt = (T)d.dt;
        dx[0] = (T)d.dx; dx[1] = (T)d.dy; dx[2] = (T)d.dz;
        cutoff = d.source_cutoff_time;
        eps_s = (T)d.eps_inv_scalar;
        mu_s = (T)d.mu_inv_scalar;
        nparts = d.complex_fields ? 2 : 1;
        for (int a = 0; a < 3; ++a) {
            periodic[a] = d.periodic[a] != 0;
            for (int q = 0; q < 2; ++q) {
                ph_fwd[a][q] = periodic[a] ? d.bloch_fwd[a][q] : (q == 0 ? 1.0 : 0.0);
                ph_rev[a][q] = periodic[a] ? d.bloch_rev[a][q] : (q == 0 ? 1.0 : 0.0);
            }
            require(nparts == 2 || (ph_fwd[a][0] == 1.0 && ph_fwd[a][1] == 0.0 && ph_rev[a][0] == 1.0 &&
                                    ph_rev[a][1] == 0.0),
                    "a Bloch phase other than 1 needs complex_fields = 1");
        }
        HIPCHK(hipGetDevice(&dev));
        st = pool_stream(dev);
        for (int p = 0; p < nparts; ++p) {
            Part &f = fp[p];
            for (int a = 0; a < 3; ++a) {
                f.E[a] = field_alloc(); f.H[a] = field_alloc();
                // D and B are state only in EXACT mode; FAST mode carries the
                // cascade on E/H and W directly
                if (mode == KH_MODE_EXACT) { f.D[a] = field_alloc(); f.B[a] = field_alloc(); }
                f.UD[a] = field_alloc(); f.UB[a] = field_alloc();
                f.WD[a] = field_alloc(); f.WB[a] = field_alloc();
            }
        }
        // sampled sigma: s[g] = profile[2g - 1] (1-based) for g = 1..N, 0 at the boundaries
        for (int a = 0; a < 3; ++a) {
            int N = g.n[a];
            for (int which = 0; which < 2; ++which) {
                const void *src = which == 0 ? d.sigma_b[a] : d.sigma_d[a];
                hassig[which][a] = src != nullptr;
                std::vector<T> h((size_t)N + 2, T(0));
                if (src) {
                    const T *ps = (const T *)src;
                    for (int i = 1; i <= N; ++i) h[i] = ps[2 * i - 2];
                }
                T *dv = (T *)dalloc(h.size() * sizeof(T));
                HIPCHK(hipMemcpyAsync(dv, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
                HIPCHK(hipStreamSynchronize(st));
                (which == 0 ? sigB : sigD)[a] = dv;
            }
        }
        // per-voxel materials, reference geometry layout -> padded grid
        for (int c = 0; c < 6; ++c) {
            const void *src = c < 3 ? d.eps_inv[c] : d.mu_inv[c - 3];
            if (!src) continue;
            T *dv = field_alloc();
            upload_material(dv, c, (const T *)src);
            (c < 3 ? epsi : mui)[c % 3] = dv;
        }
        // conductivity: per-voxel sigma and the C stage arrays
        // (src/Fields.jl needs_C: conductivity and PML arrays present)
        for (int c = 0; c < 6; ++c) {
            if (!d.sigma_mat[c]) continue;
            msig[c] = field_alloc();
            upload_material(msig[c], c, (const T *)d.sigma_mat[c], T(0));
            grow_box(cbox[c < 3 ? 1 : 0], c, (const T *)d.sigma_mat[c]);
            for (int p = 0; p < nparts; ++p) (c < 3 ? fp[p].CD : fp[p].CB)[c % 3] = field_alloc();
        }
        if (epsi[0]) rebuild_lut(1);
        if (mui[0]) rebuild_lut(0);
        partials = (double *)dalloc(NPART * 64 * sizeof(double));
        clk = (int64_t *)dalloc(256);
        graph_steps = env_int("KH_GRAPH_STEPS", 16);
        HIPCHK(hipStreamSynchronize(st));
    }

    // reference geometry layout (component voxel count, no ghosts) -> the
    // padded grid, cells 1..N; `fill` elsewhere
    void upload_material(T *dv, int c, const T *ps, T fill = T(1)) {
        int cnt[3];
        comp_count(g.n, c, cnt);
        std::vector<T> h((size_t)g.total - g.front, fill);
        for (int k = 1; k <= g.n[2]; ++k)
            for (int j = 1; j <= g.n[1]; ++j)
                for (int i = 1; i <= g.n[0]; ++i)
                    h[g.lin(i, j, k)] = ps[(size_t)(i - 1) + (size_t)cnt[0] * ((size_t)(j - 1) + (size_t)cnt[1] * (k - 1))];
        HIPCHK(hipMemcpyAsync(dv, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    // widen box b (lo[3], hi[3]) to the cells 1..N where ps (geometry layout)
    // is nonzero
    void grow_box(int *b, int c, const T *ps) const {
        int cnt[3];
        comp_count(g.n, c, cnt);
        for (int k = 1; k <= g.n[2]; ++k)
            for (int j = 1; j <= g.n[1]; ++j)
                for (int i = 1; i <= g.n[0]; ++i) {
                    if (ps[(size_t)(i - 1) + (size_t)cnt[0] * ((size_t)(j - 1) + (size_t)cnt[1] * (k - 1))] == T(0))
                        continue;
                    const int q[3] = {i, j, k};
                    for (int a = 0; a < 3; ++a) {
                        if (b[3 + a] < b[a]) { b[a] = q[a]; b[3 + a] = q[a]; }
                        b[a] = std::min(b[a], q[a]);
                        b[3 + a] = std::max(b[3 + a], q[a]);
                    }
                }
    }

    void set_eps_inv(int comp, const void *a) override {
        require(comp >= 0 && comp < 3, "component must be 0..2");
        drop_checkpoints();
        require(a != nullptr, "eps_inv array is null");
        require(epsi[comp] != nullptr, "the context has no per-voxel permittivity (create it with eps_inv arrays)");
        HIPCHK(hipStreamSynchronize(st));
        upload_material(epsi[comp], comp, (const T *)a);
        const bool was = use_lut[1];
        rebuild_lut(1);
        if (was != use_lut[1]) drop_graph();  // captured launches name the kernel variant
    }

    // Table form of group e's per-voxel material (all three components):
    // when the padded arrays hold at most 256 distinct values, a byte per
    // cell indexes a table of them.  KH_LUT=0 keeps the float arrays.
    void rebuild_lut(int e) {
        T *const *src = e ? epsi : mui;
        use_lut[e] = false;
        if (env_int("KH_LUT", 1) == 0) return;
        const size_t n = (size_t)(g.total - g.front);
        std::vector<T> vals;
        std::vector<std::vector<uint8_t>> idx(3, std::vector<uint8_t>(n));
        std::vector<T> h(n);
        for (int a = 0; a < 3; ++a) {
            HIPCHK(hipMemcpyAsync(h.data(), src[a], n * sizeof(T), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            T last = T(0);
            int li = -1;
            for (size_t i = 0; i < n; ++i) {
                const T v = h[i];
                if (li < 0 || !(v == last) || std::signbit(v) != std::signbit(last)) {
                    li = -1;
                    for (size_t k = 0; k < vals.size(); ++k)
                        if (vals[k] == v && std::signbit(vals[k]) == std::signbit(v)) { li = (int)k; break; }
                    if (li < 0) {
                        if (vals.size() == 256) return;  // too many values: float arrays
                        vals.push_back(v);
                        li = (int)vals.size() - 1;
                    }
                    last = v;
                }
                idx[a][i] = (uint8_t)li;
            }
        }
        if (!lut[e]) lut[e] = (T *)dalloc(256 * sizeof(T));
        HIPCHK(hipMemcpyAsync(lut[e], vals.data(), vals.size() * sizeof(T), hipMemcpyHostToDevice, st));
        for (int a = 0; a < 3; ++a) {
            uint8_t *&m = mid[3 * (1 - e) + a];  // mid[0..2] eps (E), mid[3..5] mu (H)
            if (!m) m = (uint8_t *)dalloc((size_t)g.total) + g.front;
            HIPCHK(hipMemcpyAsync(m, idx[a].data(), n, hipMemcpyHostToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        use_lut[e] = true;
    }

    ~Ctx() override {
        if (st) (void)hipStreamSynchronize(st);
        drop_graph();
        reap_graphs(true);
        for (auto &e : evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        for (auto &e : pool) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        for (auto &b : owned) pool_release(dev, b.p, b.bytes);
        for (void *p : ckpt) (void)hipFree(p);
        if (st) pool_stream_release(dev, st);
        if (st_adj) {
            (void)hipStreamSynchronize(st_adj);
            pool_stream_release(dev, st_adj);
        }
        for (auto &r : adj_ev)
            for (auto &e : r)
                if (e) (void)hipEventDestroy(e);
    }

    void add_source(int comp, const int32_t *s, const int32_t *d, const void *amp, int32_t *id) override {
        require(comp >= 0 && comp < 6, "component out of range");
        require(!stepped, "sources must be added before the first step", KH_E_STATE);
        int cnt[3];
        comp_count(g.n, comp, cnt);
        for (int a = 0; a < 3; ++a) {
            require(d[a] >= 1, "source extent must be positive");
            require(s[a] >= 1 && s[a] + d[a] - 1 <= cnt[a] + 1,
                    "source grid volume outside the component grid");
        }
        Source o;
        o.comp = comp;
        for (int a = 0; a < 3; ++a) { o.s[a] = s[a]; o.d[a] = d[a]; }
        size_t n = (size_t)d[0] * d[1] * d[2];
        o.amp = (T *)dalloc(2 * n * sizeof(T));
        HIPCHK(hipMemcpyAsync(o.amp, amp, 2 * n * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        srcs.push_back(std::move(o));
        ++src_version;
        *id = (int32_t)srcs.size() - 1;
    }

    void set_waveform(int id, int64_t n, const double *v) override {
        require(id >= 0 && id < (int)srcs.size(), "bad source id");
        require(n >= 0 && (n == 0 || v), "bad waveform");
        Source &o = srcs[id];
        // entries of steps already run that change make a re-run from a
        // checkpoint differ from the run: the checkpoints go
        const int64_t ran = std::min<int64_t>(ts, std::min<int64_t>(n, o.nwave));
        if (ran > 0 && std::memcmp(o.host.data(), v, (size_t)(2 * ran) * sizeof(double)) != 0) drop_checkpoints();
        if (n < o.nwave && n < ts) drop_checkpoints();
        o.host.assign(v, v + 2 * n);
        if (n > o.cap) {
            // a new table: captured step graphs hold the old pointer
            int64_t cap = std::max<int64_t>(n, 2 * o.cap);
            HIPCHK(hipStreamSynchronize(st));
            o.wave = (double *)dalloc((size_t)(2 * cap) * sizeof(double));
            o.cap = cap;
            drop_graph();
        }
        if (n > 0) {
            HIPCHK(hipMemcpyAsync(o.wave, v, (size_t)(2 * n) * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        o.nwave = n;
        ++src_version;
        drop_graph();  // nwave is a launch argument
    }

    void add_dft(int comp, const int32_t *s, const int32_t *d, int nf, const double *f, int dec, int32_t *id) override {
        // a monitor accumulates from timestep 0 like the others: the graph
        // period, the adjoint's seeding and kh_flux all assume so
        require(ts == 0, "DFT monitors must be added at timestep 0 (before stepping or after kh_reset)", KH_E_STATE);
        require(comp >= 0 && comp < 12, "component out of range");
        require(nf >= 1 && nf <= 4096, "n_freq must be in 1..4096");
        require(dec >= 1, "decimation must be >= 1");
        int cnt[3];
        comp_count(g.n, comp % 6, cnt);  // D, B sit where E, H do
        // the reference's array of the component spans indices 0..count+1; the
        // device grid holds 0..N+1, and an index past it (the ghost beyond a
        // staggered component's extra plane, never written) reads as zero
        for (int a = 0; a < 3; ++a) {
            require(d[a] >= 1, "monitor extent must be positive");
            require(s[a] >= 0 && s[a] + d[a] - 1 <= cnt[a] + 1, "monitor grid volume outside the grid");
        }
        Dft m;
        m.comp = comp;
        for (int a = 0; a < 3; ++a) { m.s[a] = s[a]; m.d[a] = d[a]; }
        m.nf = nf;
        m.dec = dec;
        m.ncell = (int64_t)d[0] * d[1] * d[2];
        m.freqs = (double *)dalloc(nf * sizeof(double));
        HIPCHK(hipMemcpyAsync(m.freqs, f, nf * sizeof(double), hipMemcpyHostToDevice, st));
        m.acc = dalloc((size_t)m.ncell * nf * 2 * (acc64 ? sizeof(double) : sizeof(T)));
        HIPCHK(hipStreamSynchronize(st));
        dfts.push_back(m);
        *id = (int32_t)dfts.size() - 1;
        if (planned) {  // after kh_reset: the DFT period may change
            drop_graph();
            plan_graph_len();
        }
    }

    // bounding box (grid indices) of the non-zero entries of an (nx, ny, nz) array
    bool nonzero_box(const T *a, int lo[3], int n[3]) const {
        int hi[3] = {0, 0, 0};
        lo[0] = lo[1] = lo[2] = 1 << 30;
        for (int k = 1; k <= g.n[2]; ++k)
            for (int j = 1; j <= g.n[1]; ++j)
                for (int i = 1; i <= g.n[0]; ++i)
                    if (a[(size_t)(i - 1) + (size_t)g.n[0] * ((size_t)(j - 1) + (size_t)g.n[1] * (k - 1))] != T(0)) {
                        lo[0] = std::min(lo[0], i); hi[0] = std::max(hi[0], i);
                        lo[1] = std::min(lo[1], j); hi[1] = std::max(hi[1], j);
                        lo[2] = std::min(lo[2], k); hi[2] = std::max(hi[2], k);
                    }
        if (hi[0] == 0) { lo[0] = lo[1] = lo[2] = 1; n[0] = n[1] = n[2] = 0; return false; }
        for (int a3 = 0; a3 < 3; ++a3) n[a3] = hi[a3] - lo[a3] + 1;
        return true;
    }

    // copy box lo..lo+n-1 of an (nx, ny, nz) array to a dense device array
    T *upload_box(const T *a, const int lo[3], const int n[3]) {
        std::vector<T> h((size_t)n[0] * n[1] * n[2]);
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i)
                    h[(size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * k)] =
                        a[(size_t)(lo[0] + i - 1) + (size_t)g.n[0] * ((size_t)(lo[1] + j - 1) + (size_t)g.n[1] * (lo[2] + k - 1))];
        T *dv = (T *)dalloc(h.size() * sizeof(T));
        HIPCHK(hipMemcpyAsync(dv, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return dv;
    }

    void set_chi3(const void *chi3) override {
        require(!stepped, "chi3 must be set before the first step", KH_E_STATE);
        require(chi3 != nullptr, "chi3 array is null");
        const T *c = (const T *)chi3;
        chi3_host.assign(c, c + (size_t)g.n[0] * g.n[1] * g.n[2]);
    }

    void add_pole(const double *c, int is_drude, const void *sigma, int32_t *id) override {
        require(!stepped, "poles must be added before the first step", KH_E_STATE);
        require(c != nullptr && sigma != nullptr, "pole coefficients or sigma array is null");
        kh::PoleDev<T> q{};
        q.g1i = (T)c[0];
        q.g1 = (T)c[1];
        // the reference's Lorentz factor is 2 - omega0_dt_sq in the backend
        // type; its Drude factor is the literal 2 (src/Kernels/Dispersive.jl:40-76)
        q.coefp = is_drude ? T(2) : T(2) - (T)c[2];
        q.drive = is_drude ? (T)c[4] : (T)c[3];
        const T *sg = (const T *)sigma;
        if (nonzero_box(sg, q.lo, q.n)) {
            q.sigma = upload_box(sg, q.lo, q.n);
            const size_t nb = (size_t)q.n[0] * q.n[1] * q.n[2] * sizeof(T);
            for (int pt = 0; pt < nparts; ++pt)
                for (int a = 0; a < 3; ++a) {
                    q.P[pt][a] = (T *)dalloc(nb);
                    q.Pp[pt][a] = (T *)dalloc(nb);
                }
            HIPCHK(hipStreamSynchronize(st));
        }
        poles.push_back(q);
        *id = (int32_t)poles.size() - 1;
    }

    // host layout: (nx+2, ny+2, nz+2), backend precision; complex fields
    // interleave (re, im)
    void pole_io(int id, int which, int comp, void *host, size_t bytes, bool rd) {
        require(id >= 0 && id < (int)poles.size(), "bad pole id");
        require(which == 0 || which == 1, "which must be 0 (P) or 1 (P_prev)");
        require(comp >= 0 && comp < 3, "polarization component must be 0..2");
        const int r[3] = {g.n[0] + 2, g.n[1] + 2, g.n[2] + 2};
        const size_t nh = (size_t)r[0] * r[1] * r[2];
        require(bytes == nh * nparts * sizeof(T), "polarization buffer size mismatch");
        kh::PoleDev<T> &q = poles[id];
        const size_t nbx = (size_t)q.n[0] * q.n[1] * q.n[2];
        T *hp = (T *)host;
        if (rd) std::fill(hp, hp + nh * nparts, T(0));
        if (nbx == 0) return;
        std::vector<T> h(nbx);
        for (int pt = 0; pt < nparts; ++pt) {
            T *dv = which ? q.Pp[pt][comp] : q.P[pt][comp];
            if (rd) {
                HIPCHK(hipMemcpyAsync(h.data(), dv, nbx * sizeof(T), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
            }
            for (int k = 0; k < q.n[2]; ++k)
                for (int j = 0; j < q.n[1]; ++j)
                    for (int i = 0; i < q.n[0]; ++i) {
                        const size_t hc = (size_t)(q.lo[0] + i) + (size_t)r[0] * ((size_t)(q.lo[1] + j) + (size_t)r[1] * (q.lo[2] + k));
                        const size_t bc = (size_t)i + (size_t)q.n[0] * ((size_t)j + (size_t)q.n[1] * k);
                        const size_t hi = nparts == 2 ? 2 * hc + pt : hc;
                        if (rd) hp[hi] = h[bc];
                        else h[bc] = hp[hi];
                    }
            if (!rd) {
                HIPCHK(hipMemcpyAsync(dv, h.data(), nbx * sizeof(T), hipMemcpyHostToDevice, st));
                HIPCHK(hipStreamSynchronize(st));
            }
        }
    }
    void read_pole(int id, int which, int comp, void *host, size_t bytes) override {
        pole_io(id, which, comp, host, bytes, true);
    }
    void write_pole(int id, int which, int comp, const void *host, size_t bytes) override {
        drop_checkpoints();
        pole_io(id, which, comp, const_cast<void *>(host), bytes, false);
        if (mat_skip) {
            mat_skip = false;
            drop_graph();  // captured material launches carry the mask
        }
    }

    // Material box = bounding box of chi3 != 0 and every pole's box; its
    // arrays are allocated on the first step.
    void plan_materials() {
        int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {0, 0, 0};
        auto take = [&](const int l[3], const int n[3]) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], l[a]);
                hi[a] = std::max(hi[a], l[a] + n[a] - 1);
            }
        };
        int clo[3], cn[3];
        bool has_chi3 = !chi3_host.empty() && nonzero_box(chi3_host.data(), clo, cn);
        if (has_chi3) take(clo, cn);
        bool has_pole = false;
        for (auto &q : poles)
            if (q.n[0] > 0) { take(q.lo, q.n); has_pole = true; }
        mat = has_chi3 || has_pole;
        if (!mat) return;
        for (int a = 0; a < 3; ++a) { mlo[a] = lo[a]; mn[a] = hi[a] - lo[a] + 1; }
        mcells = (int64_t)mn[0] * mn[1] * mn[2];
        if (has_chi3) chi3_dev = upload_box(chi3_host.data(), mlo, mn);
        chi3_host.clear();
        chi3_host.shrink_to_fit();
        const size_t nb = (size_t)mcells * sizeof(T);
        if (has_pole) {
            for (int pt = 0; pt < nparts; ++pt)
                for (int a = 0; a < 3; ++a) {
                    fpd[0][pt][a] = (T *)dalloc(nb);
                    if (mode == KH_MODE_FAST) fpd[1][pt][a] = (T *)dalloc(nb);
                }
            poles_dev = (kh::PoleDev<T> *)dalloc(poles.size() * sizeof(kh::PoleDev<T>));
            HIPCHK(hipMemcpyAsync(poles_dev, poles.data(), poles.size() * sizeof(kh::PoleDev<T>),
                                  hipMemcpyHostToDevice, st));
        }
        if (has_chi3 && mode == KH_MODE_FAST)
            for (int pt = 0; pt < nparts; ++pt)
                for (int a = 0; a < 3; ++a) delta[pt][a] = (T *)dalloc(nb);
        mat_active = (uint8_t *)dalloc((size_t)mcells);
        kh::MatArgs<T> q = mat_args();
        hipLaunchKernelGGL(kh::material_mask_kernel<T>, dim3((unsigned)((mcells + 255) / 256)), dim3(256), 0, st, q,
                           mat_active);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }

    // cells of the material box the pass updates (null: all); dropped for
    // good once a polarization array is written from the host, which can put
    // a non-zero P where its sigma is zero
    uint8_t *mat_active = nullptr;
    bool mat_skip = true;

    void launch_material() {
        kh::MatArgs<T> q = mat_args();
        const dim3 grid((unsigned)((mcells + 255) / 256)), blk(256);
        void (*kfn)(kh::MatArgs<T>);
        if (nparts == 2) kfn = mode == KH_MODE_FAST ? kh::material_kernel<T, true, 2> : kh::material_kernel<T, false, 2>;
        else kfn = mode == KH_MODE_FAST ? kh::material_kernel<T, true, 1> : kh::material_kernel<T, false, 1>;
        launch(5, kfn, grid, blk, 0, q);
    }

    kh::MatArgs<T> mat_args() const {
        kh::MatArgs<T> q{};
        for (int a = 0; a < 3; ++a) {
            for (int pt = 0; pt < nparts; ++pt) {
                q.E[pt][a] = fp[pt].E[a];
                q.W[pt][a] = fp[pt].WD[a];
                q.delta[pt][a] = delta[pt][a];
                for (int b = 0; b < 2; ++b) q.fpd[b][pt][a] = fpd[b][pt][a];
            }
            q.minv[a] = epsi[a];
            q.sig[a] = sigD[a];
            q.lo[a] = mlo[a];
            q.n[a] = mn[a];
            q.N[a] = g.n[a];
            q.per[a] = periodic[a] ? 1 : 0;
            for (int k = 0; k < 2; ++k) {
                q.fwd[a][k] = ph_fwd[a][k];
                q.rev[a][k] = ph_rev[a][k];
            }
        }
        q.minv_s = eps_s;
        q.chi3 = chi3_dev;
        q.sy = g.sy();
        q.sz = g.sz();
        q.npole = fpd[0][0][0] ? (int)poles.size() : 0;
        q.poles = poles_dev;
        q.clk = clk;
        q.koff = koff;
        q.active = mat_skip ? mat_active : nullptr;
        return q;
    }

    double round_time(int64_t n) const { return (double)((T)n * dt); }  // src/Simulation.jl:22

    // ------------------------------------------------------------ launch plans
    // Per half step (0 = H, 1 = E): the inner box (sigma zero on all axes), the
    // tile list of the full cascade kernel, and launch shapes.  Built once, on
    // the first step.
    struct Plan {
        int lo[3], hi[3];
        bool inner = false;
        kh::InnerArgs<T> ia{};
        int inner_blocks = 0;
        int nslab = 0;          // EXACT: cascade-kernel boxes (the whole grid)
        kh::Slab slab[kh::MAX_SLAB];
        int cblocks = 0;        // logical blocks of the cascade kernel
        int npslab = 0;         // FAST: PML slabs around the inner box
        kh::PSlab pslab[kh::MAX_SLAB];
        int pblocks = 0;        // logical blocks of the PML kernel
    };
    Plan plan[2];
    bool planned = false;
    bool nt = false;           // nontemporal access for single-use streams (inner kernels)
    bool src_in_pml = true;    // small FAST sources inside the PML launch (KH_SRC_LAUNCH=1: never)
    bool last_active = false;  // sources injected on the previous step

    static int env_int(const char *name, int def) {
        const char *v = getenv(name);
        return v && *v ? atoi(v) : def;
    }

    static void zero_run(const std::vector<T> &s, int N, int &lo, int &hi) {
        // longest run of zero samples among grid indices 1..N
        lo = 1; hi = 0;
        int best = 0;
        for (int i = 1; i <= N;) {
            if (s[i] != T(0)) { ++i; continue; }
            int j = i;
            while (j + 1 <= N && s[j + 1] == T(0)) ++j;
            if (j - i + 1 > best) { best = j - i + 1; lo = i; hi = j; }
            i = j + 1;
        }
    }

    void add_pslab(Plan &P, int x0, int x1, int y0, int y1, int z0, int z1) {
        // x range widened to whole pairs starting at an odd index
        if (x0 > x1 || y0 > y1 || z0 > z1) return;
        kh::PSlab &s = P.pslab[P.npslab++];
        int xs = (x0 % 2 == 1) ? x0 : x0 - 1;
        int xe = ((x1 - xs + 1) % 2 == 0) ? x1 : x1 + 1;
        s.lo[0] = xs; s.lo[1] = y0; s.lo[2] = z0;
        s.n[0] = xe - xs + 1; s.n[1] = y1 - y0 + 1; s.n[2] = z1 - z0 + 1;
        s.xlo = x0;
        s.xhi = x1;
        s.first = P.pblocks;
        int64_t pairs = (int64_t)(s.n[0] / 2) * s.n[1] * s.n[2];
        P.pblocks += (int)((pairs + kh::CBLK - 1) / kh::CBLK);
    }

    void add_slab(Plan &P, int x0, int x1, int y0, int y1, int z0, int z1) {
        if (x0 > x1 || y0 > y1 || z0 > z1) return;
        kh::Slab &s = P.slab[P.nslab++];
        s.lo[0] = x0; s.lo[1] = y0; s.lo[2] = z0;
        s.n[0] = x1 - x0 + 1; s.n[1] = y1 - y0 + 1; s.n[2] = z1 - z0 + 1;
        s.first = P.cblocks;
        int64_t cells = (int64_t)s.n[0] * s.n[1] * s.n[2];
        P.cblocks += (int)((cells + kh::CBLK - 1) / kh::CBLK);
    }

    void make_plans() {
        const int zb_inner = std::max(1, env_int("KH_ZB_INNER", 1));
        nt = env_int("KH_NT", 1) != 0;
        src_in_pml = env_int("KH_SRC_LAUNCH", 0) == 0;
        for (int e = 0; e < 2; ++e) {
            Plan &P = plan[e];
            // sampled sigma back on the host
            bool any_empty = false;
            for (int a = 0; a < 3; ++a) {
                std::vector<T> h((size_t)g.n[a] + 2);
                // all transfers go through the context stream: it is non-blocking, so
                // the legacy null stream would not be ordered after its memsets
                HIPCHK(hipMemcpyAsync(h.data(), e ? sigD[a] : sigB[a], h.size() * sizeof(T), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                zero_run(h, g.n[a], P.lo[a], P.hi[a]);
                any_empty |= P.lo[a] > P.hi[a];
            }
            // FAST mode runs the PML kernel over the whole grid (its sigma = 0
            // branch is the inner update): one launch, full 128-cell rows for
            // the x slabs, measured 1.3x faster per step than an inner-box
            // launch plus six slab launches.  The inner kernel serves grids
            // without PML (inner box = whole grid), or KH_INNER=1 forces the
            // split.
            bool whole = !any_empty;
            for (int a = 0; a < 3; ++a) whole = whole && P.lo[a] == 1 && P.hi[a] == g.n[a];
            // a group with conductivity needs the C stage of the PML kernel
            const bool cond = (e ? msig[0] || msig[1] || msig[2] : msig[3] || msig[4] || msig[5]);
            P.inner = (mode == KH_MODE_FAST) && !any_empty && !cond && (whole || env_int("KH_INNER", 0) != 0);
            if (!P.inner)
                for (int a = 0; a < 3; ++a) { P.lo[a] = 1; P.hi[a] = 0; }
            // cascade boxes: the six disjoint slabs around the inner box
            // (z slabs span the grid, y slabs the inner z range, x slabs the
            // inner y and z ranges), or the whole grid without an inner box
            P.nslab = P.cblocks = P.npslab = P.pblocks = 0;
            const int *N = g.n;
            if (mode == KH_MODE_EXACT) {
                add_slab(P, 1, N[0], 1, N[1], 1, N[2]);
            } else if (!P.inner) {
                add_pslab(P, 1, N[0], 1, N[1], 1, N[2]);
            } else {
                add_pslab(P, 1, N[0], 1, N[1], 1, P.lo[2] - 1);
                add_pslab(P, 1, N[0], 1, N[1], P.hi[2] + 1, N[2]);
                add_pslab(P, 1, N[0], 1, P.lo[1] - 1, P.lo[2], P.hi[2]);
                add_pslab(P, 1, N[0], P.hi[1] + 1, N[1], P.lo[2], P.hi[2]);
                add_pslab(P, 1, P.lo[0] - 1, P.lo[1], P.hi[1], P.lo[2], P.hi[2]);
                add_pslab(P, P.hi[0] + 1, N[0], P.lo[1], P.hi[1], P.lo[2], P.hi[2]);
            }
            if (P.inner) {
                kh::InnerArgs<T> &ia = P.ia;
                for (int a = 0; a < 3; ++a) {
                    // field pointers are set per part at launch
                    ia.minv[a] = e ? epsi[a] : mui[a];
                    ia.inv[a] = T(1) / dx[a];
                    ia.lo[a] = P.lo[a];
                    ia.hi[a] = P.hi[a];
                }
                ia.minv_s = e ? eps_s : mu_s;
                ia.dt = dt;
                ia.sy = g.sy();
                ia.sz = g.sz();
                ia.zb = std::min(zb_inner, P.hi[2] - P.lo[2] + 1);
                int kx0 = (P.lo[0] - 1) / (2 * kh::TX), kx1 = (P.hi[0] - 1) / (2 * kh::TX);
                ia.x0 = 1 + 2 * kh::TX * kx0;
                ia.tiles[0] = kx1 - kx0 + 1;
                ia.tiles[1] = (P.hi[1] - P.lo[1] + 1 + kh::TY - 1) / kh::TY;
                ia.tiles[2] = (P.hi[2] - P.lo[2] + 1 + ia.zb - 1) / ia.zb;
                int nt = ia.tiles[0] * ia.tiles[1] * ia.tiles[2];
                ia.per_xcd = kh::xcd_per(nt);
                P.inner_blocks = ia.per_xcd * kh::NXCD;
                bool pervox = (e ? epsi[0] : mui[0]) != nullptr;
                for (int a = 1; a < 3; ++a)
                    require(((e ? epsi[a] : mui[a]) != nullptr) == pervox,
                            "per-voxel materials must be given for all three components or none");
            }
        }
        plan_materials();
        plan_xfold();
        planned = true;
    }

    void plan_xfold() {
        const int N = g.n[0];
        bool ok = mode == KH_MODE_FAST && periodic[0] && !plan[0].inner && !plan[1].inner && !fields_written;
        if (ok && mat) ok = mlo[0] > 1 && mlo[0] + mn[0] - 1 < N;
        // E/H monitors on an x ghost plane form its value from the opposite
        // plane (dft_kernel); FAST D/B monitors rebuild from stage arrays
        // that have no ghost copies, so they keep the ghost fill
        for (auto &m : dfts) ok = ok && (m.comp < 6 || (m.s[0] > 0 && m.s[0] + m.d[0] - 1 < N + 1));
        if (ok != xfold) drop_graph();
        if (xfold && !ok) refresh_xghosts();
        xfold = ok;
    }

    // fill the x ghost planes of E and H as the per-half-step wrap would
    void refresh_xghosts() {
        if (!xghost_stale) return;
        launch_wrap_axis(0, 0);
        launch_wrap_axis(1, 0);
        xghost_stale = false;
    }

    kh::CascadeArgs<T> cascade_args(int e, int part) const {
        const Plan &P = plan[e];
        const Part &f = fp[part];
        kh::CascadeArgs<T> p{};
        for (int a = 0; a < 3; ++a) {
            p.A[a] = e ? f.E[a] : f.H[a];
            p.F[a] = e ? f.H[a] : f.E[a];
            p.Tq[a] = e ? f.D[a] : f.B[a];
            p.U[a] = e ? f.UD[a] : f.UB[a];
            p.W[a] = e ? f.WD[a] : f.WB[a];
            p.sig[a] = e ? sigD[a] : sigB[a];
            p.minv[a] = e ? epsi[a] : mui[a];
            p.inv[a] = T(1) / dx[a];
            p.msig[a] = msig[e ? a : 3 + a];
            p.C[a] = e ? f.CD[a] : f.CB[a];
            // no next-axis profile: the reference allocates no U for the
            // component (src/Fields.jl needs_U)
            p.noU[a] = !hassig[e][(a + 1) % 3];
        }
        p.minv_s = e ? eps_s : mu_s;
        p.dt = dt;
        p.sy = g.sy();
        p.sz = g.sz();
        p.nslab = P.nslab;
        for (int s = 0; s < P.nslab; ++s) p.slab[s] = P.slab[s];
        p.nblock = P.cblocks;
        p.per_xcd = kh::xcd_per(P.cblocks);
        if (e && fpd[0][0][0]) {
            for (int a = 0; a < 3; ++a) {
                p.fpd[a] = fpd[0][part][a];
                p.ulo[a] = mlo[a];
                p.un[a] = mn[a];
            }
        }
        return p;
    }

    // FAST sources the PML launch of this half step applies itself (empty:
    // none, or source_fix_kernel launches)
    std::vector<kh::FixSrc<T>> pml_srcs;

    // all of the sources together in at most 1024 wave-wide row segments of
    // the PML launch (a wave: 64 lanes x 2 cells)
    static bool pml_sources_small(const std::vector<kh::FixSrc<T>> &list) {
        int64_t rows = 0;
        for (const auto &o : list) rows += (int64_t)o.d[1] * o.d[2] * ((o.d[0] + 127) / 128 + 1);
        return rows <= 1024;
    }

    // Sources precomputed per block of steps: a half step whose sources (at
    // most MAX_SRC) are all on one component with pairwise disjoint boxes has
    // their corrections for the whole block computed by one launch at the
    // block's start (source_dw_kernel, from the device clock), and the PML
    // launch adds them in registers before its stores: no per-step source
    // launch and no wait on the waveform tables inside the update.
    struct PrePlan {
        bool ok = false;
        int comp = 0, n = 0, cap = 0;  // cap: steps a buffer row set holds
        int64_t cells = 0;
        kh::PreBox box[kh::MAX_SRC];
        T *buf = nullptr;
        size_t bytes = 0;
    };
    PrePlan pre[2];
    uint64_t pre_version = 0;
    bool pre_ready[2] = {false, false};  // the current block's corrections are computed

    void plan_pre() {
        plan_dfte();
        if (pre_version == src_version) return;
        pre_version = src_version;
        for (int e = 0; e < 2; ++e) {
            PrePlan &P = pre[e];
            P.ok = false;
            // (real fields without conductivity: the kernels' PRE variants)
            const bool cond = e ? (msig[0] || msig[1] || msig[2]) : (msig[3] || msig[4] || msig[5]);
            if (mode != KH_MODE_FAST || !src_in_pml || nparts != 1 || cond || plan[e].pblocks == 0 || plan[e].inner)
                continue;
            const std::vector<kh::FixSrc<T>> list = fix_sources(e == 1);
            if (list.empty() || list.size() > (size_t)kh::MAX_SRC) continue;
            bool one = true;
            for (size_t i = 0; i < list.size() && one; ++i) {
                one = list[i].own == list[0].own;
                for (size_t j = 0; j < i && one; ++j) one = !boxes_meet(list[i], list[j]);
            }
            if (!one) continue;
            // small sources stay with the end-of-lane form (no extra launch)
            if (pml_sources_small(list)) continue;
            int64_t cells = 0;
            for (size_t i = 0; i < list.size(); ++i) {
                kh::PreBox &b = P.box[i];
                for (int k = 0; k < 3; ++k) { b.s[k] = list[i].s[k]; b.d[k] = list[i].d[k]; }
                b.off = cells;
                cells += (int64_t)b.d[0] * b.d[1] * b.d[2];
            }
            // rows for a main graph block, a short block or a checkpoint
            // interval, within 64 MiB
            int want = std::max(std::max(graph_len, graph_steps), 1);
            if (ckpt_every > 0 && ckpt_every <= 1024) want = std::max(want, ckpt_every);
            const int64_t fit = ((int64_t)64 << 20) / (cells * (int64_t)sizeof(T));
            const int cap = (int)std::min<int64_t>(want, fit);
            if (cap < 1) continue;
            const size_t bytes = (size_t)cells * cap * sizeof(T);
            if (bytes > P.bytes) {
                // a new buffer: graphs captured with the old one are dropped
                drop_graph();
                P.buf = (T *)dalloc(bytes);
                P.bytes = bytes;
            }
            P.comp = list[0].own;
            P.n = (int)list.size();
            P.cells = cells;
            P.cap = cap;
            P.ok = true;
        }
    }

    // corrections for the next len steps (block step k: device clock + k);
    // plan_pre() has run since the sources last changed (it allocates, so
    // never inside a capture)
    void launch_pre(int len) {
        launch_dfte_weights(len);
        for (int e = 0; e < 2; ++e) {
            pre_ready[e] = false;
            const PrePlan &P = pre[e];
            if (!P.ok || !(active || last_active) || len > P.cap) continue;
            const std::vector<kh::FixSrc<T>> list = fix_sources(e == 1);
            kh::PreSrcArgs<T> a{};
            a.n = P.n;
            for (int i = 0; i < P.n; ++i) {
                a.box[i] = P.box[i];
                a.amp[i] = list[i].amp;
                a.wave[i] = list[i].wave;
                a.nwave[i] = list[i].nwave;
            }
            a.prev = list[0].prev;
            a.minv = list[0].minv;
            a.minv_s = e ? eps_s : mu_s;
            a.sig_prev = list[0].sig_prev;
            a.dt = dt;
            a.cutoff = cutoff;
            a.clk = clk;
            a.nsteps = len;
            a.cells = P.cells;
            a.n0 = g.n[0]; a.n1 = g.n[1]; a.n2 = g.n[2];
            a.sy = g.sy();
            a.sz = g.sz();
            a.dw = P.buf;
            const int64_t total = (int64_t)len * P.cells;
            hipLaunchKernelGGL(kh::source_dw_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a);
            HIPCHK(hipGetLastError());
            pre_ready[e] = true;
        }
    }
    void end_pre() {
        pre_ready[0] = pre_ready[1] = false;
        dfte_ready = false;
    }

    // Small E monitors whose DFT update the E launch applies as it stores
    // (FAST mode, whole-grid PML launch, real fields, float64 accumulators,
    // no material pass or periodic axis, at most MAX_DFTE monitors of at most
    // 8192 cell-frequencies inside cells 1..N, no E source writing a monitor
    // cell of its component; KH_DFT_LAUNCH=1: never): their weights for a
    // block of steps come from one dft_weights_kernel launch at the block's
    // start, and dft_kernel runs only for the other monitors.
    std::vector<int> dfte;  // indices into dfts
    std::vector<int> dfte_off;
    int dfte_stride = 0, dfte_cap = 0;
    double *dfte_buf = nullptr;
    size_t dfte_bytes = 0;
    bool dfte_ready = false;

    void plan_dfte() {
        std::vector<int> sel, off;
        int stride = 0;
        bool ok = mode == KH_MODE_FAST && nparts == 1 && acc64 && !mat && plan[1].pblocks > 0 && !plan[1].inner &&
                  env_int("KH_DFT_LAUNCH", 0) == 0;
        for (int a = 0; a < 3; ++a) ok = ok && !periodic[a] && !msig[a];
        for (size_t i = 0; ok && i < dfts.size() && (int)sel.size() < kh::MAX_DFTE; ++i) {
            const Dft &m = dfts[i];
            if (m.comp >= 3 || m.ncell * m.nf > 8192) continue;
            bool inside = true;
            for (int k = 0; k < 3; ++k) inside = inside && m.d[k] > 0 && m.s[k] >= 1 && m.s[k] + m.d[k] - 1 <= g.n[k];
            if (!inside) continue;
            bool hit = false;
            for (auto &o : srcs) {
                if (o.comp != m.comp) continue;
                bool meet = true;
                for (int k = 0; k < 3 && meet; ++k) meet = o.s[k] < m.s[k] + m.d[k] && m.s[k] < o.s[k] + o.d[k];
                hit = hit || meet;
            }
            if (hit) continue;
            sel.push_back((int)i);
            off.push_back(stride);
            stride += 1 + 2 * m.nf;
        }
        int cap = 0;
        if (!sel.empty()) {
            int want = std::max(std::max(graph_len, graph_steps), 1);
            if (ckpt_every > 0 && ckpt_every <= 1024) want = std::max(want, ckpt_every);
            const int64_t fit = ((int64_t)64 << 20) / ((int64_t)stride * (int64_t)sizeof(double));
            cap = (int)std::min<int64_t>(want, fit);
            if (cap < 1) sel.clear();
        }
        if (sel != dfte || off != dfte_off || stride != dfte_stride) drop_graph();  // launch arguments change
        dfte = sel;
        dfte_off = off;
        dfte_stride = stride;
        dfte_cap = sel.empty() ? 0 : cap;
        const size_t bytes = (size_t)dfte_cap * stride * sizeof(double);
        if (bytes > dfte_bytes) {
            drop_graph();
            dfte_buf = (double *)dalloc(bytes);
            dfte_bytes = bytes;
        }
    }

    // weights of the E update's monitors for the next len steps (row k:
    // device clock + k)
    void launch_dfte_weights(int len) {
        dfte_ready = false;
        if (dfte.empty() || len > dfte_cap) return;
        kh::DftW<T> w{};
        w.nmon = (int)dfte.size();
        for (size_t i = 0; i < dfte.size(); ++i) {
            const Dft &m = dfts[dfte[i]];
            w.freqs[i] = m.freqs;
            w.nf[i] = m.nf;
            w.dec[i] = m.dec;
            w.woff[i] = dfte_off[i];
            w.elec[i] = 1;
        }
        w.clk = clk;
        w.koff0 = 0;
        w.len = len;
        w.stride = dfte_stride;
        w.dt = dt;
        w.out = dfte_buf;
        const int64_t n = (int64_t)len * dfte_stride;
        hipLaunchKernelGGL((kh::dft_weights_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w);
        HIPCHK(hipGetLastError());
        dfte_ready = true;
    }

    // design-box snapshot written by the E update of the steps launched
    // while set (the adjoint's re-run; null: none)
    T *snap_out[3] = {nullptr, nullptr, nullptr};
    int snap_lo[3][3] = {}, snap_n[3][3] = {};

    // the E update can write the snapshot of these boxes itself: FAST mode
    // with the whole-grid PML launch, real fields, no conductivity, and no E
    // source writing a box cell of its component (sources applied after the
    // update's stores would change the value after the snapshot was taken);
    // KH_SNAP_LAUNCH=1 keeps the separate box_copy_kernel
    bool snap_fusable(const int lo[3][3], const int n[3][3]) const {
        if (env_int("KH_SNAP_LAUNCH", 0) != 0) return false;
        if (mode != KH_MODE_FAST || nparts != 1 || plan[1].pblocks == 0 || plan[1].inner || plan[1].cblocks)
            return false;
        for (int a = 0; a < 3; ++a)
            if (fp[0].CD[a]) return false;
        for (auto &o : srcs) {
            if (o.comp >= 3) continue;
            const int a = o.comp;
            bool meet = n[a][0] > 0 && n[a][1] > 0 && n[a][2] > 0;
            for (int k = 0; k < 3 && meet; ++k)
                meet = o.s[k] < lo[a][k] + n[a][k] && lo[a][k] < o.s[k] + o.d[k];
            if (meet) return false;
        }
        return true;
    }

    void launch_half(int e) {
        const bool electric = e == 1;
        // complex fields without conductivity: one PML launch updates both parts
        bool fuse = nparts == 2 && plan[e].pblocks > 0;
        for (int a = 0; a < 3; ++a) fuse = fuse && !(e ? fp[0].CD[a] : fp[0].CB[a]);
        // FAST mode sources (real part only: S += Re(f A),
        // src/Sources/Sources.jl:346-357) while sources are or were on the
        // previous step; the switch is permanent and the kernels gate
        // themselves by the device timestep.  Three forms, the same
        // operations on the same values:
        //  * corrections precomputed for the block (plan_pre: real fields,
        //    one component, disjoint boxes, not small), added by the PML
        //    launch before its stores;
        //  * small sources (points, lines: all of them together in at most
        //    1024 wave-wide row segments of the PML launch) applied by that
        //    launch as its lanes finish, in registration order; a lane in a
        //    source box waits for the waveform tables after its stores, which
        //    costs less than a launch only while few waves do it (a 192 x 160
        //    plane made the E update 14% slower than a separate launch);
        //  * otherwise source_fix_kernel after the update.
        std::vector<kh::FixSrc<T>> list;
        if (mode == KH_MODE_FAST && (active || last_active)) list = fix_sources(electric);
        pml_srcs.clear();
        if (pre_ready[e]) {
            list.clear();  // the PML launch adds the precomputed corrections
        } else if (src_in_pml && !list.empty() && plan[e].pblocks > 0 && !plan[e].inner &&
                   list.size() <= (size_t)kh::MAX_SRC && pml_sources_small(list)) {
            pml_srcs.swap(list);
        }
        for (int part = 0; part < nparts; ++part) launch_half_part(e, part, fuse);
        pml_srcs.clear();
        // Sources in separate launches share a launch unless two of them write
        // the same cell of one component: then they go to successive
        // launches, in registration order.
        if (!list.empty()) {
            size_t q0 = 0;
            while (q0 < list.size()) {
                kh::FixArgs<T> q{};
                int blocks = 0;
                size_t q1 = q0;
                for (; q1 < list.size() && q1 - q0 < (size_t)kh::MAX_SRC; ++q1) {
                    bool clash = false;
                    for (size_t r = q0; r < q1; ++r) clash = clash || boxes_meet(list[r], list[q1]);
                    if (clash) break;
                    q.src[q.nsrc] = list[q1];
                    q.src[q.nsrc].first = blocks;
                    const auto &x = list[q1];
                    blocks += (int)(((int64_t)x.d[0] * x.d[1] * x.d[2] + 255) / 256);
                    ++q.nsrc;
                }
                q0 = q1;
                q.minv_s = electric ? eps_s : mu_s;
                q.dt = dt;
                q.cutoff = cutoff;
                q.clk = clk;
                q.koff = koff;
                q.n0 = g.n[0]; q.n1 = g.n[1]; q.n2 = g.n[2];
                q.sy = g.sy();
                q.sz = g.sz();
                if (blocks == 0) continue;
                hipLaunchKernelGGL((kh::source_fix_kernel<T>), dim3(blocks), dim3(256), 0, st, q);
                HIPCHK(hipGetLastError());
            }
        }
        launch_wrap(e);
    }

    // EXACT: the sources of each half step as a device list read by the
    // cascade kernels (any number of them), brought up to date before the
    // step's launches are issued or captured
    kh::SrcL<T> *srcl_dev[2] = {nullptr, nullptr};
    int srcl_n[2] = {0, 0}, srcl_cap[2] = {0, 0};
    uint64_t src_version = 1, srcl_version = 0;

    void sync_source_lists() {
        if (mode != KH_MODE_EXACT || srcl_version == src_version) return;
        for (int e = 0; e < 2; ++e) {
            std::vector<kh::SrcL<T>> l;
            for (auto &o : srcs) {
                if ((o.comp < 3) != (e == 1)) continue;
                kh::SrcL<T> x{};
                x.comp = o.comp % 3;
                for (int a = 0; a < 3; ++a) { x.s[a] = o.s[a]; x.d[a] = o.d[a]; }
                x.amp = o.amp;
                x.wave = o.wave;
                x.nwave = o.nwave;
                l.push_back(x);
            }
            if ((int)l.size() > srcl_cap[e]) {
                // a new list: captured graphs hold the old pointer
                drop_graph();
                srcl_cap[e] = std::max<int>((int)l.size(), 2 * srcl_cap[e]);
                srcl_dev[e] = (kh::SrcL<T> *)dalloc((size_t)srcl_cap[e] * sizeof(kh::SrcL<T>));
            }
            if (!l.empty())
                HIPCHK(hipMemcpyAsync(srcl_dev[e], l.data(), l.size() * sizeof(kh::SrcL<T>), hipMemcpyHostToDevice, st));
            if ((int)l.size() != srcl_n[e]) drop_graph();  // the count is a launch argument
            srcl_n[e] = (int)l.size();
        }
        HIPCHK(hipStreamSynchronize(st));
        srcl_version = src_version;
    }

    std::vector<kh::FixSrc<T>> fix_sources(bool electric) const {
        std::vector<kh::FixSrc<T>> out;
        const Part &f = fp[0];
        for (auto &o : srcs) {
            if ((o.comp < 3) != electric) continue;
            kh::FixSrc<T> x{};
            const int a = o.comp % 3, prv = (a + 2) % 3;
            x.A = electric ? f.E[a] : f.H[a];
            x.W = electric ? f.WD[a] : f.WB[a];
            x.minv = electric ? epsi[a] : mui[a];
            x.sig_own = electric ? sigD[a] : sigB[a];
            x.sig_prev = electric ? sigD[prv] : sigB[prv];
            x.own = a;
            x.prev = prv;
            x.amp = o.amp;
            for (int k = 0; k < 3; ++k) { x.s[k] = o.s[k]; x.d[k] = o.d[k]; }
            x.wave = o.wave;
            x.nwave = o.nwave;
            out.push_back(x);
        }
        return out;
    }

    // two sources that write a common cell of one component
    static bool boxes_meet(const kh::FixSrc<T> &a, const kh::FixSrc<T> &b) {
        if (a.A != b.A) return false;
        for (int k = 0; k < 3; ++k)
            if (a.s[k] + a.d[k] <= b.s[k] || b.s[k] + b.d[k] <= a.s[k]) return false;
        return true;
    }

    // periodic / Bloch ghost fill of the group just updated, axis by axis
    void launch_wrap(int e) {
        for (int ax = 0; ax < 3; ++ax) {
            if (!periodic[ax]) continue;
            if (ax == 0 && xfold) {
                xghost_stale = true;
                continue;
            }
            launch_wrap_axis(e, ax);
        }
    }

    void launch_wrap_axis(int e, int ax) {
        {
            kh::WrapArgs<T> w{};
            for (int pt = 0; pt < nparts; ++pt)
                for (int a = 0; a < 3; ++a) w.F[pt][a] = e ? fp[pt].E[a] : fp[pt].H[a];
            w.nparts = nparts;
            w.ax = ax;
            for (int a = 0; a < 3; ++a) w.N[a] = g.n[a];
            w.sy = g.sy();
            w.sz = g.sz();
            for (int k = 0; k < 2; ++k) {
                w.fwd[k] = ph_fwd[ax][k];
                w.rev[k] = ph_rev[ax][k];
            }
            w.unit = ph_fwd[ax][0] == 1.0 && ph_fwd[ax][1] == 0.0 && ph_rev[ax][0] == 1.0 && ph_rev[ax][1] == 0.0;
            const int64_t n = (int64_t)g.n[(ax + 1) % 3] * g.n[(ax + 2) % 3] * 6;
            require(n < (1LL << 31) - 256, "periodic face too large");
            hipLaunchKernelGGL((kh::wrap_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w);
            HIPCHK(hipGetLastError());
        }
    }

    void launch_half_part(int e, int part, bool fuse) {
        const Plan &P = plan[e];
        const bool electric = e == 1;
        const Part &f = fp[part];
        dim3 blk(kh::TX, kh::TY, 1);
        if (P.inner) {
            kh::InnerArgs<T> ia = P.ia;
            for (int a = 0; a < 3; ++a) {
                ia.A[a] = electric ? f.E[a] : f.H[a];
                ia.F[a] = electric ? f.H[a] : f.E[a];
            }
            const bool pervox = ia.minv[0] != nullptr;
            void (*kfn)(kh::InnerArgs<T>);
            if (electric) {
                if (pervox) kfn = nt ? kh::inner_e_kernel<T, true, true> : kh::inner_e_kernel<T, true, false>;
                else kfn = nt ? kh::inner_e_kernel<T, false, true> : kh::inner_e_kernel<T, false, false>;
            } else {
                if (pervox) kfn = nt ? kh::inner_h_kernel<T, true, true> : kh::inner_h_kernel<T, true, false>;
                else kfn = nt ? kh::inner_h_kernel<T, false, true> : kh::inner_h_kernel<T, false, false>;
            }
            launch(2 * e, kfn, dim3(P.inner_blocks), blk, 0, ia);
        }
        if (P.cblocks) {
            kh::CascadeArgs<T> p = cascade_args(e, part);
            // sources add Re(f A) to the real part only
            p.nsrc = part == 0 ? srcl_n[e] : 0;
            p.src = srcl_dev[e];
            p.clk = clk;
            p.koff = koff;
            p.cutoff = cutoff;
            const dim3 cgrid(p.per_xcd * kh::NXCD), cblk(kh::CBLK);
            launch(2 * e + 1, electric ? kh::cascade_e_kernel<T> : kh::cascade_h_kernel<T>, cgrid, cblk, 0, p);
        }
        if (P.pblocks && !(fuse && part == 1)) {
            kh::PmlArgs<T> q{};
            for (int k = 0; k < (fuse ? 2 : 1); ++k) {
                const Part &fk = fp[fuse ? k : part];
                for (int a = 0; a < 3; ++a) {
                    q.f[k].A[a] = electric ? fk.E[a] : fk.H[a];
                    q.f[k].F[a] = electric ? fk.H[a] : fk.E[a];
                    q.f[k].U[a] = electric ? fk.UD[a] : fk.UB[a];
                    q.f[k].W[a] = electric ? fk.WD[a] : fk.WB[a];
                }
            }
            for (int a = 0; a < 3; ++a) {
                q.sig[a] = electric ? sigD[a] : sigB[a];
                q.minv[a] = electric ? epsi[a] : mui[a];
                q.inv[a] = T(1) / dx[a];
                q.msig[a] = msig[electric ? a : 3 + a];
                q.C[a] = electric ? f.CD[a] : f.CB[a];
                q.clo[a] = cbox[e][a];
                q.chi[a] = cbox[e][3 + a];
            }
            const bool cond = q.C[0] || q.C[1] || q.C[2];
            for (int a = 0; a < 3; ++a) q.mid[a] = mid[(electric ? 0 : 3) + a];
            q.lut = lut[e];
            q.perx = xfold ? 1 : 0;
            q.nxg = g.n[0];
            if (xfold) {
                // the H update reads E's ghost N+1 (forward phase), the E update
                // H's ghost 0 (reverse phase), as the wrap fills them
                const double *ph = electric ? ph_rev[0] : ph_fwd[0];
                const bool cplx = nparts == 2 && !(ph_fwd[0][0] == 1.0 && ph_fwd[0][1] == 0.0 &&
                                                   ph_rev[0][0] == 1.0 && ph_rev[0][1] == 0.0);
                for (int k = 0; k < (fuse ? 2 : 1); ++k) {
                    const int pk = fuse ? k : part;
                    q.f[k].wa = cplx ? ph[0] : 1.0;
                    q.f[k].wb = cplx ? (pk == 0 ? -ph[1] : ph[1]) : 0.0;
                    for (int a = 0; a < 3; ++a)
                        q.f[k].Fo[a] = cplx ? (electric ? fp[1 - pk].H[a] : fp[1 - pk].E[a]) : nullptr;
                }
            }
            q.minv_s = electric ? eps_s : mu_s;
            q.dt = dt;
            q.sy = g.sy();
            q.sz = g.sz();
            // sources go to the real part (f[0] of a two-part launch)
            if (part == 0) {
                q.nsrc = (int)pml_srcs.size();
                for (size_t k = 0; k < pml_srcs.size(); ++k) q.src[k] = pml_srcs[k];
                if (pre_ready[e]) {
                    const PrePlan &PP = pre[e];
                    q.pre_n = PP.n;
                    q.pre_comp = PP.comp;
                    for (int k = 0; k < PP.n; ++k) q.pre[k] = PP.box[k];
                    q.pre_dw = PP.buf;
                    q.pre_cells = PP.cells;
                }
            }
            const bool snap = electric && part == 0 && !fuse && !cond && (snap_out[0] || snap_out[1] || snap_out[2]);
            const bool dfte_on = electric && part == 0 && !fuse && !cond && !snap && dfte_ready;
            if (dfte_on) {
                q.ndfte = (int)dfte.size();
                for (size_t i = 0; i < dfte.size(); ++i) {
                    const Dft &m = dfts[dfte[i]];
                    kh::DfteMon &x = q.dfte[i];
                    x.acc = (double *)m.acc;
                    x.comp = m.comp;
                    for (int k = 0; k < 3; ++k) { x.s[k] = m.s[k]; x.d[k] = m.d[k]; }
                    x.nf = m.nf;
                    x.woff = dfte_off[i];
                    x.ncell = m.ncell;
                }
                q.dftw = dfte_buf + (size_t)koff * dfte_stride;
            }
            if (snap)
                for (int a = 0; a < 3; ++a) {
                    q.snap[a] = snap_out[a];
                    for (int k = 0; k < 3; ++k) {
                        q.snap_lo[a][k] = snap_lo[a][k];
                        q.snap_n[a][k] = snap_n[a][k];
                    }
                }
            q.clk = clk;
            q.koff = koff;
            q.cutoff = cutoff;
            q.nslab = P.npslab;
            for (int s = 0; s < P.npslab; ++s) q.slab[s] = P.pslab[s];
            q.nblock = P.pblocks;
            q.per_xcd = kh::xcd_per(P.pblocks);
            const bool pervox = q.minv[0] != nullptr;
            const dim3 pgrid(q.per_xcd * kh::NXCD), pblk(kh::CBLK);
            const bool lut_on = pervox && use_lut[e];
            // template flags: per-voxel material, conductivity C stage,
            // material bytes + table, both parts of a complex field
            void (*kfn)(kh::PmlArgs<T>);
            if (dfte_on) {
                if (q.pre_n > 0) kfn = lut_on ? kh::pml_e_kernel<T, true, false, true, false, true, false, true>
                                     : pervox ? kh::pml_e_kernel<T, true, false, false, false, true, false, true>
                                              : kh::pml_e_kernel<T, false, false, false, false, true, false, true>;
                else kfn = lut_on ? kh::pml_e_kernel<T, true, false, true, false, false, false, true>
                         : pervox ? kh::pml_e_kernel<T, true, false, false, false, false, false, true>
                                  : kh::pml_e_kernel<T, false, false, false, false, false, false, true>;
            } else if (snap) {
                if (q.pre_n > 0) kfn = lut_on ? kh::pml_e_kernel<T, true, false, true, false, true, true>
                                     : pervox ? kh::pml_e_kernel<T, true, false, false, false, true, true>
                                              : kh::pml_e_kernel<T, false, false, false, false, true, true>;
                else kfn = lut_on ? kh::pml_e_kernel<T, true, false, true, false, false, true>
                         : pervox ? kh::pml_e_kernel<T, true, false, false, false, false, true>
                                  : kh::pml_e_kernel<T, false, false, false, false, false, true>;
            } else if (fuse) {
                if (electric) kfn = lut_on ? kh::pml_e_kernel<T, true, false, true, true>
                                  : pervox ? kh::pml_e_kernel<T, true, false, false, true>
                                           : kh::pml_e_kernel<T, false, false, false, true>;
                else kfn = lut_on ? kh::pml_h_kernel<T, true, false, true, true>
                         : pervox ? kh::pml_h_kernel<T, true, false, false, true>
                                  : kh::pml_h_kernel<T, false, false, false, true>;
            } else if (q.pre_n > 0) {
                if (electric) kfn = lut_on ? kh::pml_e_kernel<T, true, false, true, false, true>
                                  : pervox ? kh::pml_e_kernel<T, true, false, false, false, true>
                                           : kh::pml_e_kernel<T, false, false, false, false, true>;
                else kfn = lut_on ? kh::pml_h_kernel<T, true, false, true, false, true>
                         : pervox ? kh::pml_h_kernel<T, true, false, false, false, true>
                                  : kh::pml_h_kernel<T, false, false, false, false, true>;
            } else if (electric) {
                if (lut_on) kfn = cond ? kh::pml_e_kernel<T, true, true, true> : kh::pml_e_kernel<T, true, false, true>;
                else if (cond) kfn = pervox ? kh::pml_e_kernel<T, true, true> : kh::pml_e_kernel<T, false, true>;
                else kfn = pervox ? kh::pml_e_kernel<T, true, false> : kh::pml_e_kernel<T, false, false>;
            } else {
                if (lut_on) kfn = cond ? kh::pml_h_kernel<T, true, true, true> : kh::pml_h_kernel<T, true, false, true>;
                else if (cond) kfn = pervox ? kh::pml_h_kernel<T, true, true> : kh::pml_h_kernel<T, false, true>;
                else kfn = pervox ? kh::pml_h_kernel<T, true, false> : kh::pml_h_kernel<T, false, false>;
            }
            launch(2 * e + 1, kfn, pgrid, pblk, 0, q);
        }
    }

    // All DFT monitors of a half step, in launches of up to MAX_DFT monitors;
    // each gates itself by its decimation against the device timestep.
    void launch_dfts(bool electric) {
        // the host knows the timestep of every launched or captured step
        // (graphs start on a multiple of every decimation): only due
        // monitors are launched; the kernel gates itself as well
        const int64_t step_ts = ts + koff;
        std::vector<const Dft *> list;
        for (size_t i = 0; i < dfts.size(); ++i) {
            const Dft &m = dfts[i];
            if (electric && dfte_ready && std::find(dfte.begin(), dfte.end(), (int)i) != dfte.end())
                continue;  // applied by the E launch
            if ((m.comp % 6 < 3) == electric && (dft_gate_device || m.dec <= 1 || step_ts % m.dec == 0))
                list.push_back(&m);
        }
        for (size_t i0 = 0; i0 < list.size(); i0 += kh::MAX_DFT) {
            kh::DftArgs<T> q{};
            int blocks = 0, nfmax = 1;
            for (size_t i = i0; i < list.size() && q.nmon < kh::MAX_DFT; ++i) {
                const Dft &m = *list[i];
                kh::DftMon<T> &x = q.mon[q.nmon++];
                x.acc = m.acc;
                const int a = m.comp % 3;
                const bool db = m.comp >= 6;  // D or B
                if (db && mode == KH_MODE_EXACT) {
                    x.f = electric ? fp[0].D[a] : fp[0].B[a];
                    x.fim = nparts == 2 ? (electric ? fp[1].D[a] : fp[1].B[a]) : nullptr;
                } else {
                    x.f = electric ? fp[0].E[a] : fp[0].H[a];
                    x.fim = nparts == 2 ? (electric ? fp[1].E[a] : fp[1].H[a]) : nullptr;
                }
                if (db && mode == KH_MODE_FAST) {
                    x.rebuild = 1;
                    x.w = electric ? fp[0].WD[a] : fp[0].WB[a];
                    x.wim = nparts == 2 ? (electric ? fp[1].WD[a] : fp[1].WB[a]) : nullptr;
                    x.minv = electric ? epsi[a] : mui[a];
                    x.minv_s = electric ? eps_s : mu_s;
                    x.sig_own = electric ? sigD[a] : sigB[a];
                    x.own = a;
                }
                x.freqs = m.freqs;
                for (int k = 0; k < 3; ++k) { x.s[k] = m.s[k]; x.d[k] = m.d[k]; }
                x.nf = m.nf;
                x.dec = m.dec;
                x.first = blocks;
                blocks += (int)((m.ncell + 255) / 256);
                nfmax = std::max(nfmax, m.nf);
            }
            q.clk = clk;
            q.koff = koff;
            q.dt = dt;
            q.electric = electric ? 1 : 0;
            q.n0 = g.n[0]; q.n1 = g.n[1]; q.n2 = g.n[2];
            q.sy = g.sy();
            q.sz = g.sz();
            q.perx = xfold ? 1 : 0;
            q.xunit = ph_fwd[0][0] == 1.0 && ph_fwd[0][1] == 0.0 && ph_rev[0][0] == 1.0 && ph_rev[0][1] == 0.0;
            for (int k = 0; k < 2; ++k) {
                q.xfwd[k] = ph_fwd[0][k];
                q.xrev[k] = ph_rev[0][k];
            }
            const size_t shm = (size_t)nfmax * 2 * sizeof(double);
            launch(4, acc64 ? kh::dft_kernel<T, double> : kh::dft_kernel<T, T>, dim3(blocks), dim3(256), shm, q);
            ++(capturing ? cap_dfts : n_dft);
        }
    }

    // One time step's launches (src/Kernels/Kernels.jl:64-87 order); every
    // time-dependent quantity comes from the device timestep.
    int koff = 0;                 // offset of the step being launched in its block
    bool dft_gate_device = false;  // decimations too long for a graph period: launch every step
    int64_t dft_period = 1;        // lcm of the DFT decimations

    void launch_step(int k) {
        koff = k;
        launch_half(0);
        launch_dfts(false);
        launch_half(1);
        if (mat) launch_material();
        launch_dfts(true);
        koff = 0;
    }

    void launch_tick(int64_t n) {
        hipLaunchKernelGGL(kh::tick_kernel, dim3(1), dim3(64), 0, st, clk, n);
        HIPCHK(hipGetLastError());
    }

    // host mirror of the source switch for the step about to run
    void begin_step() {
        if (active && round_time(ts) > cutoff) active = false;  // src/Kernels/Kernels.jl:27-35
    }
    void end_step() {
        last_active = active;
        ++ts;
    }

    // ---------------------------------------------------------------- graphs
    // Steps are captured as hipGraphs and replayed (the reference captures its
    // H and E updates after the sources switch off,
    // src/Kernels/Kernels.jl:99-145; here sources and monitors are inside the
    // graph too, gated by the device timestep).  The host decides which DFT
    // monitors are due at each captured position from the timestep the
    // block starts at, so a graph is valid for replays that start at the same
    // phase modulo the DFT period:
    //   * the main graph: graph_len steps (graph_steps rounded up to the
    //     period), started on multiples of the period (phase 0 every replay);
    //   * short graphs: graph_steps steps from any phase, for runs shorter
    //     than graph_len or not aligned to the period; the last few phases
    //     used are kept.
    // Re-captured when a launch argument changes (a new waveform table, the
    // FAST source correction dropping out after the switch-off).
    struct StepGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        int len = 0;
        int64_t phase = -1;
        bool src = false;      // the captured steps include the FAST source correction
        int64_t dfts = 0;      // DFT launches captured
        uint64_t used = 0;     // last use (short-graph replacement)
    };
    // A replaced graph may still have launches queued on the stream: it is
    // destroyed once an event recorded behind them has completed.
    struct DeadGraph {
        hipGraph_t g;
        hipGraphExec_t x;
        hipEvent_t ev;
    };
    std::vector<DeadGraph> dead_graphs;

    void retire_graph(StepGraph &sg) {
        if (sg.x || sg.g) {
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
                hipEventRecord(ev, st) == hipSuccess) {
                dead_graphs.push_back({sg.g, sg.x, ev});
            } else {
                // no event: wait for the stream instead
                (void)hipGetLastError();
                if (ev) (void)hipEventDestroy(ev);
                (void)hipStreamSynchronize(st);
                if (sg.x) (void)hipGraphExecDestroy(sg.x);
                if (sg.g) (void)hipGraphDestroy(sg.g);
            }
        }
        sg.x = nullptr;
        sg.g = nullptr;
        sg.phase = -1;
        reap_graphs(false);
    }

    // destroy the retired graphs whose launches have completed (all: wait)
    void reap_graphs(bool wait) {
        for (auto it = dead_graphs.begin(); it != dead_graphs.end();) {
            const hipError_t e = wait ? hipEventSynchronize(it->ev) : hipEventQuery(it->ev);
            if (e == hipErrorNotReady) {
                ++it;
                continue;
            }
            (void)hipGetLastError();
            if (it->x) (void)hipGraphExecDestroy(it->x);
            if (it->g) (void)hipGraphDestroy(it->g);
            (void)hipEventDestroy(it->ev);
            it = dead_graphs.erase(it);
        }
    }
    static constexpr int N_SHORT = 4;
    int graph_steps = 16;
    StepGraph main_graph;
    StepGraph short_graph[N_SHORT];
    uint64_t graph_clock = 0;
    bool graph_failed = false;
    // launch counts since the last kh_get_launch_counts
    int64_t n_replays = 0, n_graph_steps = 0, n_eager = 0, n_dft = 0;
    bool capturing = false;
    int64_t cap_dfts = 0;

    void drop_graph() {
        retire_graph(main_graph);
        for (auto &s : short_graph) retire_graph(s);
    }

    bool capture(StepGraph &sg, int len) {
        retire_graph(sg);
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        bool ok = true;
        capturing = true;
        cap_dfts = 0;
        try {
            launch_pre(len);
            for (int i = 0; i < len; ++i) launch_step(i);
            launch_tick(len);
        } catch (const KhError &) {
            ok = false;
        }
        end_pre();
        capturing = false;
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (!ok || e != hipSuccess || !gr) {
            if (gr) (void)hipGraphDestroy(gr);
            (void)hipGetLastError();
            return false;
        }
        sg.g = gr;
        if (hipGraphInstantiate(&sg.x, sg.g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipGraphDestroy(sg.g);
            sg.g = nullptr;
            sg.x = nullptr;
            return false;
        }
        sg.len = len;
        sg.phase = ts % dft_period;
        sg.src = active || last_active;
        sg.dfts = cap_dfts;
        return true;
    }

    int graph_len = 16;  // steps per main graph: graph_steps rounded up to the DFT period

    void plan_graph_len() {
        dft_period = 1;
        for (auto &m : dfts) {
            const int64_t d = std::max(1, m.dec);
            int64_t a = dft_period, b = d;
            while (b) { const int64_t t = a % b; a = b; b = t; }
            dft_period = dft_period / a * d;
            if (dft_period > 256) break;
        }
        dft_gate_device = dft_period > 256;
        if (dft_gate_device) dft_period = 1;
        graph_len = graph_steps <= 0 ? 0 : (int)(((graph_steps + dft_period - 1) / dft_period) * dft_period);
    }

    // The graph the next `left` steps from the current host state replay,
    // or nullptr for an eager step.  begin_step() has run for the first step.
    StepGraph *pick_graph(int64_t left) {
        if (graph_len <= 0 || timing || graph_failed) return nullptr;
        const auto fits = [&](int len) { return ckpt_every <= 0 || ts % ckpt_every + len <= ckpt_every; };
        if (left >= graph_len && ts % dft_period == 0 && fits(graph_len)) return &main_graph;
        if (graph_len > graph_steps && left >= graph_steps && fits(graph_steps)) {
            const int64_t ph = ts % dft_period;
            StepGraph *hit = nullptr, *old = &short_graph[0];
            for (auto &s : short_graph) {
                if (s.phase == ph) hit = &s;
                if (s.used < old->used) old = &s;
            }
            return hit ? hit : old;
        }
        return nullptr;
    }

    // Capture (if needed) the graph for the current host state; false: eager.
    bool ready(StepGraph *sg, int64_t left) {
        const int len = sg == &main_graph ? graph_len : graph_steps;
        const bool src_now = active || last_active;
        if (!sg->x || sg->src != src_now || sg->len != len || sg->phase != ts % dft_period) {
            if (!capture(*sg, len)) {
                graph_failed = true;
                return false;
            }
        }
        (void)left;
        sg->used = ++graph_clock;
        return true;
    }

    void step(int64_t n) override {
        require(n >= 0, "n_steps must be >= 0");
        if (!planned) {
            make_plans();
            plan_graph_len();
        }
        stepped = true;
        sync_source_lists();
        plan_pre();
        if (xfold && n > 0) xghost_stale = true;
        int64_t done = 0;
        while (done < n) {
            if (ckpt_every > 0 && ts % ckpt_every == 0) save_checkpoint();
            begin_step();
            StepGraph *sg = pick_graph(n - done);
            if (sg && ready(sg, n - done)) {
                HIPCHK(hipGraphLaunch(sg->x, st));
                ++n_replays;
                n_graph_steps += sg->len;
                n_dft += sg->dfts;
                end_step();
                for (int i = 1; i < sg->len; ++i) {
                    begin_step();
                    end_step();
                }
                done += sg->len;
                continue;
            }
            launch_pre(1);
            launch_step(0);
            end_pre();
            launch_tick(1);
            end_step();
            ++n_eager;
            ++done;
        }
    }

    // Capture the graphs that step(n) would replay from the current state
    // without running anything, so that the first replays of a timed run do
    // not pay for capture.
    void prepare_graphs(int64_t n) override {
        require(n >= 0, "n_steps must be >= 0");
        if (!planned) {
            make_plans();
            plan_graph_len();
        }
        sync_source_lists();
        plan_pre();
        const int64_t ts0 = ts;
        const bool a0 = active, l0 = last_active;
        int64_t done = 0;
        while (done < n) {
            begin_step();
            StepGraph *sg = pick_graph(n - done);
            const int len = sg && ready(sg, n - done) ? sg->len : 1;
            end_step();
            for (int i = 1; i < len; ++i) {
                begin_step();
                end_step();
            }
            done += len;
        }
        ts = ts0;
        active = a0;
        last_active = l0;
    }

    // steps of the main graph (the block a run of steps is best cut into)
    int graph_block() override {
        if (!planned) {
            make_plans();
            plan_graph_len();
        }
        return graph_len > 0 ? graph_len : 1;
    }

    void launch_counts(int64_t *out) override {
        out[0] = n_replays;
        out[1] = n_graph_steps;
        out[2] = n_eager;
        out[3] = n_dft;
        n_replays = n_graph_steps = n_eager = n_dft = 0;
    }

    void set_graph(int steps) override {
        require(steps >= 0 && steps <= 4096, "graph steps must be 0..4096");
        graph_steps = steps;
        graph_failed = false;
        drop_graph();
        if (planned) plan_graph_len();
    }

    void sync() override { HIPCHK(hipStreamSynchronize(st)); }
    bool idle() override {
        hipError_t e = hipStreamQuery(st);
        if (e == hipErrorNotReady) return false;
        HIPCHK(e);
        return true;
    }
    int64_t timestep() const override { return ts; }
    bool sources_active() const override { return active; }

    T *field_ptr(int comp, int part = 0) const { return comp < 3 ? fp[part].E[comp] : fp[part].H[comp - 3]; }

    // Reference layout: (component voxel count + 2) per axis, column-major;
    // grid indices 0..N+1 of every axis live on the device grid (an index past
    // N+1 on a staggered axis reads as zero and is not written).  Complex
    // fields interleave (re, im).
    void read_field(int comp, void *host, size_t bytes) override {
        require(comp >= 0 && comp < 6, "component out of range");
        refresh_xghosts();
        int cnt[3];
        comp_count(g.n, comp, cnt);
        int r[3] = {cnt[0] + 2, cnt[1] + 2, cnt[2] + 2};
        size_t nh = (size_t)r[0] * r[1] * r[2];
        require(bytes == nh * nparts * sizeof(T), "field buffer size mismatch");
        std::vector<T> h((size_t)g.total - g.front);
        T *o = (T *)host;
        for (int part = 0; part < nparts; ++part) {
            HIPCHK(hipMemcpyAsync(h.data(), field_ptr(comp, part), h.size() * sizeof(T), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < r[2]; ++k)
                for (int j = 0; j < r[1]; ++j)
                    for (int i = 0; i < r[0]; ++i) {
                        T v = T(0);
                        if (i <= g.n[0] + 1 && j <= g.n[1] + 1 && k <= g.n[2] + 1) v = h[g.lin(i, j, k)];
                        const size_t hi = (size_t)i + (size_t)r[0] * ((size_t)j + (size_t)r[1] * k);
                        o[nparts == 2 ? 2 * hi + part : hi] = v;
                    }
        }
    }

    void write_field(int comp, const void *host, size_t bytes) override {
        require(comp >= 0 && comp < 6, "component out of range");
        drop_checkpoints();
        // written ghosts are read as written: the x ghost fill goes back to
        // every half step (after bringing the other ghosts up to date)
        fields_written = true;
        if (planned) plan_xfold();
        int cnt[3];
        comp_count(g.n, comp, cnt);
        int r[3] = {cnt[0] + 2, cnt[1] + 2, cnt[2] + 2};
        size_t nh = (size_t)r[0] * r[1] * r[2];
        require(bytes == nh * nparts * sizeof(T), "field buffer size mismatch");
        // D/B and the PML auxiliaries are left untouched
        std::vector<T> h((size_t)g.total - g.front, T(0));
        const T *in = (const T *)host;
        const int m[3] = {std::min(r[0] - 1, g.n[0] + 1), std::min(r[1] - 1, g.n[1] + 1),
                          std::min(r[2] - 1, g.n[2] + 1)};
        for (int part = 0; part < nparts; ++part) {
            for (int k = 0; k <= m[2]; ++k)
                for (int j = 0; j <= m[1]; ++j)
                    for (int i = 0; i <= m[0]; ++i) {
                        const size_t hi = (size_t)i + (size_t)r[0] * ((size_t)j + (size_t)r[1] * k);
                        h[g.lin(i, j, k)] = in[nparts == 2 ? 2 * hi + part : hi];
                    }
            HIPCHK(hipMemcpyAsync(field_ptr(comp, part), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
            // a written E is taken as the linear field: no chi3 deficit carried over
            if (comp < 3 && delta[part][comp])
                HIPCHK(hipMemsetAsync(delta[part][comp], 0, (size_t)mcells * sizeof(T), st));
        }
        HIPCHK(hipStreamSynchronize(st));
    }

    void read_dft(int id, double *host, size_t bytes) override {
        require(id >= 0 && id < (int)dfts.size(), "bad dft id");
        Dft &m = dfts[id];
        size_t n = (size_t)m.ncell * m.nf * 2;
        require(bytes == n * sizeof(double), "dft buffer size mismatch");
        if (acc64) {
            HIPCHK(hipMemcpyAsync(host, m.acc, n * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        } else {
            std::vector<T> h(n);
            HIPCHK(hipMemcpyAsync(h.data(), m.acc, n * sizeof(T), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (size_t i = 0; i < n; ++i) host[i] = (double)h[i];
        }
    }

    void flux(int e1, int e2, int h1, int h2, int ax1, double dA, double *out) override {
        int ids[4] = {e1, e2, h1, h2};
        for (int q = 0; q < 4; ++q) require(ids[q] >= 0 && ids[q] < (int)dfts.size(), "bad dft id");
        require(ax1 >= 1 && ax1 <= 3, "normal axis must be 1..3");
        int nf = dfts[e1].nf;
        for (int q = 0; q < 4; ++q) require(dfts[ids[q]].nf == nf, "flux monitors differ in n_freq");
        int ax = ax1 - 1, t0 = (ax + 1) % 3, t1 = (ax + 2) % 3;
        if (t0 > t1) std::swap(t0, t1);
        int n1 = 1 << 30, n2 = 1 << 30;
        kh::FluxMon fm[4];
        for (int q = 0; q < 4; ++q) {
            Dft &m = dfts[ids[q]];
            fm[q].acc = m.acc;
            for (int a = 0; a < 3; ++a) fm[q].d[a] = m.d[a];
            n1 = std::min(n1, m.d[t0]);
            n2 = std::min(n2, m.d[t1]);
        }
        require(nf <= NPART * 64, "too many frequencies");
        if (acc64)
            hipLaunchKernelGGL(kh::flux_kernel<double>, dim3(nf), dim3(256), 0, st, fm[0], fm[1], fm[2], fm[3],
                               ax, t0, t1, n1, n2, dA, partials);
        else
            hipLaunchKernelGGL(kh::flux_kernel<T>, dim3(nf), dim3(256), 0, st, fm[0], fm[1], fm[2], fm[3],
                               ax, t0, t1, n1, n2, dA, partials);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, partials, nf * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    double *mode_buf = nullptr;   // mode profiles of the last overlap (grown on demand)
    size_t mode_cap = 0;

    void mode_overlap(int e1, int e2, int h1, int h2, int ax1, int n1, int n2, const double *mode, double dA,
                      double *out) override {
        int ids[4] = {e1, e2, h1, h2};
        for (int q = 0; q < 4; ++q) require(ids[q] >= 0 && ids[q] < (int)dfts.size(), "bad dft id");
        require(ax1 >= 1 && ax1 <= 3, "normal axis must be 1..3");
        const int nf = dfts[e1].nf;
        for (int q = 0; q < 4; ++q) require(dfts[ids[q]].nf == nf, "mode monitor DFTs differ in n_freq");
        int ax = ax1 - 1, t0 = (ax + 1) % 3, t1 = (ax + 2) % 3;
        if (t0 > t1) std::swap(t0, t1);
        kh::FluxMon fm[4];
        for (int q = 0; q < 4; ++q) {
            Dft &m = dfts[ids[q]];
            fm[q].acc = m.acc;
            for (int a = 0; a < 3; ++a) fm[q].d[a] = m.d[a];
            require(n1 >= 1 && n2 >= 1 && n1 <= m.d[t0] && n2 <= m.d[t1], "mode plane larger than a DFT plane");
        }
        require(nf <= NPART * 16, "too many frequencies");
        const size_t bytes = (size_t)4 * nf * n1 * n2 * 2 * sizeof(double);
        if (bytes > mode_cap) {
            mode_buf = (double *)dalloc(bytes);
            mode_cap = bytes;
        }
        HIPCHK(hipMemcpyAsync(mode_buf, mode, bytes, hipMemcpyHostToDevice, st));
        if (acc64)
            hipLaunchKernelGGL(kh::mode_overlap_kernel<double>, dim3(nf), dim3(256), 0, st, fm[0], fm[1], fm[2],
                               fm[3], ax, t0, t1, n1, n2, nf, (const double *)mode_buf, dA, partials);
        else
            hipLaunchKernelGGL(kh::mode_overlap_kernel<T>, dim3(nf), dim3(256), 0, st, fm[0], fm[1], fm[2], fm[3],
                               ax, t0, t1, n1, n2, nf, (const double *)mode_buf, dA, partials);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, partials, (size_t)4 * nf * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    // growable device buffer owned by the context (the old one is freed)
    void *regrow(void *old, size_t &cap, size_t bytes) {
        if (bytes <= cap) return old;
        if (old) {
            HIPCHK(hipStreamSynchronize(st));
            for (auto it = owned.begin(); it != owned.end(); ++it)
                if (it->p == old) {
                    pool_release(dev, it->p, it->bytes);
                    owned.erase(it);
                    break;
                }
        }
        cap = bytes;
        return dalloc(bytes);
    }

    double *n2f_in = nullptr, *n2f_part = nullptr, *n2f_out = nullptr;  // near-to-far staging
    size_t n2f_in_cap = 0, n2f_part_cap = 0, n2f_out_cap = 0;

    double *n2f_surf = nullptr;   // gathered surface: values (nf x N x 8), positions (N x 12)
    size_t n2f_surf_cap = 0;

    void near2far(const int32_t *ids, int ax1, double ns, const double *base, double d1, double d2, double dA,
                  int n1, int n2, const double *freqs, int nf, double eps, double mu, const double *obs, int n_obs,
                  double *out) override {
        for (int q = 0; q < 4; ++q) require(ids[q] >= 0 && ids[q] < (int)dfts.size(), "bad dft id");
        require(ax1 >= 1 && ax1 <= 3, "normal axis must be 1..3");
        require(n_obs >= 1 && nf >= 1, "need at least one observation point and one frequency");
        require(eps > 0 && mu > 0 && std::isfinite(eps) && std::isfinite(mu), "medium eps and mu must be > 0");
        require(std::isfinite(d1) && std::isfinite(d2) && std::isfinite(dA) && std::isfinite(ns),
                "near-to-far needs a 3D grid (finite spacings)");
        for (int q = 0; q < 4; ++q) require(dfts[ids[q]].nf == nf, "near-to-far DFTs differ from n_freq");
        const int ax = ax1 - 1;
        const int t0 = ax == 0 ? 1 : 0, t1 = ax == 2 ? 1 : 2;
        kh::FluxMon fm[4];
        for (int q = 0; q < 4; ++q) {
            Dft &m = dfts[ids[q]];
            fm[q].acc = m.acc;
            for (int a = 0; a < 3; ++a) fm[q].d[a] = m.d[a];
            require(n1 >= 1 && n2 >= 1 && n1 <= m.d[t0] && n2 <= m.d[t1], "surface larger than a DFT plane");
            require(m.d[ax] <= 2, "a near-to-far DFT spans more than two planes across the normal");
        }
        const int64_t N = (int64_t)n1 * n2;
        const int64_t m_out = (int64_t)nf * 6 * n_obs;
        require(N < (1LL << 31) && m_out < (1LL << 31) && nf <= 65535, "near-to-far problem too large");
        kh::N2FGeom geo;
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a) {
                geo.base[q][a] = base[3 * q + a];
                require(std::isfinite(geo.base[q][a]), "non-finite base position");
            }
        geo.d1 = d1;
        geo.d2 = d2;
        geo.dA = dA;
        geo.ns = ns;
        geo.eps = eps;
        geo.mu = mu;
        geo.Z = std::sqrt(mu / eps);
        const bool unit = eps == 1.0 && mu == 1.0;
        // block: one thread per observation point (64, 128 or 256 of them)
        const int bs = n_obs <= 64 ? 64 : (n_obs <= 128 ? 128 : 256);
        const int64_t blocks_xy = (int64_t)((n_obs + bs - 1) / bs) * nf;
        const int64_t tiles = (N + bs - 1) / bs;
        // split the samples until the grid holds ~32 waves per SIMD (about ten
        // rounds of resident blocks, so the last round costs little), keeping the
        // partial fields within 512 MB
        const int64_t want = (32768 / (bs / 64) + blocks_xy - 1) / blocks_xy;
        const int64_t cap_mem = std::max<int64_t>(1, (512LL << 20) / (m_out * 16));
        int splits = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(want, tiles), std::min<int64_t>(cap_mem, 65535)));
        const int64_t tiles_per = (tiles + splits - 1) / splits;
        splits = (int)((tiles + tiles_per - 1) / tiles_per);
        const int chunk = (int)(tiles_per * bs);
        const size_t in_bytes = ((size_t)3 * n_obs + nf) * sizeof(double);
        n2f_in = (double *)regrow(n2f_in, n2f_in_cap, in_bytes);
        HIPCHK(hipMemcpyAsync(n2f_in, obs, (size_t)3 * n_obs * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(n2f_in + 3 * n_obs, freqs, (size_t)nf * sizeof(double), hipMemcpyHostToDevice, st));
        const size_t out_bytes = (size_t)m_out * 2 * sizeof(double);
        n2f_out = (double *)regrow(n2f_out, n2f_out_cap, out_bytes);
        double *dst = n2f_out;
        if (splits > 1) {
            n2f_part = (double *)regrow(n2f_part, n2f_part_cap, out_bytes * splits);
            dst = n2f_part;
        }
        // gathered surface for up to nfc frequencies at a time (<= 512 MB)
        const int nfc = (int)std::max<int64_t>(1, std::min<int64_t>(nf, ((512LL << 20) / (8 * (int64_t)N * 8)) - 2));
        n2f_surf = (double *)regrow(n2f_surf, n2f_surf_cap, ((size_t)nfc * 8 + 12) * N * sizeof(double));
        double *vals = n2f_surf, *pos = n2f_surf + (size_t)nfc * 8 * N;
        const double *o_dev = n2f_in, *f_dev = n2f_in + 3 * n_obs;
        // EXACT contexts evaluate green3d! in the reference's operation order;
        // FAST ones use green3d_fast (same phase argument, amplitudes to ~1 ulp)
        const bool fast = mode == KH_MODE_FAST;
        auto go = [&](auto ax_c, auto unit_c, auto fast_c) {
            constexpr int A = decltype(ax_c)::value;
            constexpr bool U = decltype(unit_c)::value, F = decltype(fast_c)::value;
            for (int k0 = 0; k0 < nf; k0 += nfc) {
                const int nk = std::min(nfc, nf - k0);
                const dim3 gg((unsigned)((N + 255) / 256), nk);
                if (acc64)
                    hipLaunchKernelGGL((kh::n2f_gather_kernel<double, A, F>), gg, dim3(256), 0, st, fm[0], fm[1],
                                       fm[2], fm[3], geo, n1, n2, k0, vals, pos);
                else
                    hipLaunchKernelGGL((kh::n2f_gather_kernel<T, A, F>), gg, dim3(256), 0, st, fm[0], fm[1], fm[2],
                                       fm[3], geo, n1, n2, k0, vals, pos);
                HIPCHK(hipGetLastError());
                dim3 grid((n_obs + bs - 1) / bs, nk, splits);
                hipLaunchKernelGGL((kh::near2far_kernel<A, U, F>), grid, dim3(bs), 0, st, (const double *)vals,
                                   (const double *)pos, geo, o_dev, f_dev, n_obs, nf, k0, (int)N, chunk, dst);
                HIPCHK(hipGetLastError());
            }
        };
        using Fl = std::false_type;
        using Tr = std::true_type;
        auto go2 = [&](auto ax_c) {
            if (unit) fast ? go(ax_c, Tr{}, Tr{}) : go(ax_c, Tr{}, Fl{});
            else fast ? go(ax_c, Fl{}, Tr{}) : go(ax_c, Fl{}, Fl{});
        };
        if (ax == 0) go2(std::integral_constant<int, 0>{});
        else if (ax == 1) go2(std::integral_constant<int, 1>{});
        else go2(std::integral_constant<int, 2>{});
        HIPCHK(hipGetLastError());
        if (splits > 1) {
            hipLaunchKernelGGL(kh::n2f_reduce_kernel, dim3((unsigned)((m_out + 255) / 256)), dim3(256), 0, st,
                               (const double *)n2f_part, splits, m_out, n2f_out);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipMemcpyAsync(out, n2f_out, out_bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    int32_t *n2f_bins = nullptr;
    double *n2f_spec = nullptr;
    size_t n2f_bins_cap = 0, n2f_spec_cap = 0;

    void near2far_spectrum(const int32_t *ids, int ax1, int n1, int n2, int k, int p1, int p2, const int32_t *bins,
                           int n_bins, double *out) override {
        for (int q = 0; q < 4; ++q) require(ids[q] >= 0 && ids[q] < (int)dfts.size(), "bad dft id");
        require(ax1 >= 1 && ax1 <= 3, "normal axis must be 1..3");
        require(n_bins >= 1, "need at least one bin");
        require(p1 >= n1 && p2 >= n2 && p1 >= 1 && p2 >= 1, "padded size smaller than the plane");
        const int ax = ax1 - 1;
        const int t0 = ax == 0 ? 1 : 0, t1 = ax == 2 ? 1 : 2;
        kh::FluxMon fm[4];
        for (int q = 0; q < 4; ++q) {
            Dft &m = dfts[ids[q]];
            require(k >= 0 && k < m.nf, "frequency index out of range");
            fm[q].acc = m.acc;
            for (int a = 0; a < 3; ++a) fm[q].d[a] = m.d[a];
            require(n1 >= 1 && n2 >= 1 && n1 <= m.d[t0] && n2 <= m.d[t1], "surface larger than a DFT plane");
            require(m.d[ax] <= 2, "a near-to-far DFT spans more than two planes across the normal");
        }
        require((int64_t)n1 * n2 < (1LL << 31) && n_bins < (1 << 30), "spectrum problem too large");
        n2f_bins = (int32_t *)regrow(n2f_bins, n2f_bins_cap, (size_t)n_bins * 2 * sizeof(int32_t));
        n2f_spec = (double *)regrow(n2f_spec, n2f_spec_cap, (size_t)n_bins * 8 * sizeof(double));
        HIPCHK(hipMemcpyAsync(n2f_bins, bins, (size_t)n_bins * 2 * sizeof(int32_t), hipMemcpyHostToDevice, st));
        auto go = [&](auto ax_c) {
            constexpr int A = decltype(ax_c)::value;
            if (acc64)
                hipLaunchKernelGGL((kh::n2f_spectrum_kernel<double, A>), dim3(n_bins), dim3(256), 0, st, fm[0],
                                   fm[1], fm[2], fm[3], n1, n2, k, p1, p2, (const int *)n2f_bins, n2f_spec);
            else
                hipLaunchKernelGGL((kh::n2f_spectrum_kernel<T, A>), dim3(n_bins), dim3(256), 0, st, fm[0], fm[1],
                                   fm[2], fm[3], n1, n2, k, p1, p2, (const int *)n2f_bins, n2f_spec);
        };
        if (ax == 0) go(std::integral_constant<int, 0>{});
        else if (ax == 1) go(std::integral_constant<int, 1>{});
        else go(std::integral_constant<int, 2>{});
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, n2f_spec, (size_t)n_bins * 8 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    // every monitor's norm, in registration order: launches into separate
    // partial slots, one copy and one synchronisation per 64 monitors
    void dft_norms(int n_out, double *out) override {
        require(n_out == (int)dfts.size(), "n_out must equal the number of DFT monitors");
        std::vector<double> part((size_t)NPART * 64);
        for (int q0 = 0; q0 < n_out; q0 += 64) {
            const int nq = std::min(64, n_out - q0);
            for (int k = 0; k < nq; ++k) {
                Dft &m = dfts[q0 + k];
                const int64_t n = m.ncell * m.nf;
                double *dst = partials + (size_t)k * NPART;
                if (acc64)
                    hipLaunchKernelGGL(kh::sumsq_kernel<double>, dim3(NPART), dim3(256), 0, st, (const double *)m.acc, n, dst);
                else
                    hipLaunchKernelGGL(kh::sumsq_kernel<T>, dim3(NPART), dim3(256), 0, st, (const T *)m.acc, n, dst);
                HIPCHK(hipGetLastError());
            }
            HIPCHK(hipMemcpyAsync(part.data(), partials, (size_t)nq * NPART * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < nq; ++k) {
                double s = 0.0;
                for (int b = 0; b < NPART; ++b) s += part[(size_t)k * NPART + b];
                out[q0 + k] = std::sqrt(s);
            }
        }
    }

    void dft_norm(int id, double *out) override {
        require(id >= -1 && id < (int)dfts.size(), "bad dft id");
        double total = 0.0;
        std::vector<double> part(NPART);
        for (int q = 0; q < (int)dfts.size(); ++q) {
            if (id >= 0 && q != id) continue;
            Dft &m = dfts[q];
            int64_t n = m.ncell * m.nf;
            if (acc64)
                hipLaunchKernelGGL(kh::sumsq_kernel<double>, dim3(NPART), dim3(256), 0, st, (const double *)m.acc, n, partials);
            else
                hipLaunchKernelGGL(kh::sumsq_kernel<T>, dim3(NPART), dim3(256), 0, st, (const T *)m.acc, n, partials);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(part.data(), partials, NPART * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            double s = 0.0;
            for (double v : part) s += v;
            total += std::sqrt(s);
        }
        *out = total;
    }

    // ------------------------------------------- checkpoints and the adjoint
    // Every ckpt_every steps of a run from timestep 0 the whole field state
    // (E, H, D, B, the PML stages) is copied to a checkpoint; the adjoint
    // re-runs each interval from its checkpoint and walks it backwards.
    int ckpt_every = 0;
    std::vector<void *> ckpt;
    struct CkHost {
        int64_t ts = -1;
        bool active = true, last_active = false, valid = false;
    };
    std::vector<CkHost> ck_host;

    std::vector<T *> state_arrays() const {
        std::vector<T *> v;
        for (int part = 0; part < nparts; ++part) {
            const Part &f = fp[part];
            for (int a = 0; a < 3; ++a) {
                T *arrs[10] = {f.E[a], f.H[a], f.D[a], f.B[a], f.UD[a], f.UB[a], f.WD[a], f.WB[a], f.CD[a], f.CB[a]};
                for (T *q : arrs)
                    if (q) v.push_back(q);
            }
        }
        return v;
    }

    // the whole field state to (save) or from a buffer of state_bytes() (one
    // array every state_stride() bytes), in one copy launch for up to
    // MAX_STATE arrays
    size_t state_stride() const { return ((size_t)g.total * sizeof(T) + 255) / 256 * 256; }
    void copy_state(void *dst, const void *src_state, bool save) {
        const auto arrs = state_arrays();
        const size_t fb = (size_t)g.total * sizeof(T), stride = state_stride();
        for (size_t i0 = 0; i0 < arrs.size(); i0 += kh::MAX_STATE) {
            kh::StateCopy c{};
            const size_t n = std::min(arrs.size() - i0, (size_t)kh::MAX_STATE);
            for (size_t j = 0; j < n; ++j) {
                char *buf = (char *)(save ? dst : const_cast<void *>(src_state)) + (i0 + j) * stride;
                void *arr = arrs[i0 + j] - g.front;
                c.src[j] = save ? (const void *)arr : (const void *)buf;
                c.dst[j] = save ? (void *)buf : arr;
            }
            c.w16 = (int64_t)(fb / 16);
            c.tail = (int)((fb % 16) / 4);
            const dim3 grid((unsigned)((c.w16 + c.tail + 255) / 256), (unsigned)n);
            hipLaunchKernelGGL(kh::state_copy_kernel, grid, dim3(256), 0, st, c);
            HIPCHK(hipGetLastError());
        }
    }

    size_t state_bytes() const { return state_arrays().size() * state_stride(); }

    void save_checkpoint() {
        const size_t k = (size_t)(ts / ckpt_every);
        while (ckpt.size() <= k) {
            void *q = nullptr;
            HIPCHK(hipMalloc(&q, state_bytes()));
            ckpt.push_back(q);
            ck_host.push_back(CkHost{});
        }
        copy_state(ckpt[k], nullptr, true);
        ck_host[k] = CkHost{ts, active, last_active, true};
    }

    // state written from outside the time step: a re-run from an earlier
    // checkpoint would no longer reproduce the run, so none is kept (the
    // checkpoint of the current timestep is taken when stepping resumes)
    void drop_checkpoints() {
        for (auto &c : ck_host) c.valid = false;
    }

    void restore_state(const void *src, const CkHost &h) {
        copy_state(nullptr, src, false);
        if (xfold) xghost_stale = true;
        // the clock is set in stream order, so the host need not wait
        hipLaunchKernelGGL(kh::set_clock_kernel, dim3(1), dim3(64), 0, st, clk, (int64_t)h.ts);
        HIPCHK(hipGetLastError());
        ts = h.ts;
        active = h.active;
        last_active = h.last_active;
    }

    void set_checkpoints(int every) override {
        require(every >= 0, "checkpoint interval must be >= 0");
        require(every == 0 || ts == 0, "enable checkpoints before the first step or after kh_reset", KH_E_STATE);
        HIPCHK(hipStreamSynchronize(st));
        for (auto &c : ck_host) c.valid = false;
        if (every == 0) {
            for (void *q : ckpt) HIPCHK(hipFree(q));
            ckpt.clear();
            ck_host.clear();
        }
        ckpt_every = every;
    }

    // adjoint state, per group (0: H, 1: E) and component: the adjoints of
    // A (E/H), T (D/B), U, W and Kbar; allocated on the first adjoint call
    T *adjA[2][3]{}, *adjT[2][3]{}, *adjU[2][3]{}, *adjW[2][3]{}, *adjK[2][3]{};
    void *adj_final = nullptr;
    double *adj_cot = nullptr, *adj_grad = nullptr;
    T *adj_seg = nullptr;
    double *adj_sw = nullptr;  // step weight table of the objective's seeds
    // the sweep's second stream (walk back of one interval while the next
    // interval is re-run), its events, and the stream the transposed half
    // steps are launched on (null: st)
    hipStream_t st_adj = nullptr;
    hipEvent_t adj_ev[2][2] = {};  // [re-run done / walk back done][buffer]
    hipStream_t ast = nullptr;
    size_t adj_final_cap = 0, adj_cot_cap = 0, adj_grad_cap = 0, adj_seg_cap = 0, adj_sw_cap = 0;

    // The seeds of the step at device timestep + koff; every monitor of the
    // group is launched and skips the steps its decimation does not sample,
    // so the launches do not depend on the step.
    void launch_adj_seed(const std::vector<const Dft *> &mons, const std::vector<size_t> &off, bool electric,
                         int64_t koff) {
        std::vector<size_t> list;
        for (size_t i = 0; i < mons.size(); ++i)
            if ((mons[i]->comp < 3) == electric) list.push_back(i);
        // Monitors of one launch update lambda without atomics, so two that
        // write a common cell of one component (overlapping objectives, or
        // one monitor listed twice) go to successive launches, in list order.
        auto meet = [](const Dft &a, const Dft &b) {
            if (a.comp % 3 != b.comp % 3) return false;
            for (int k = 0; k < 3; ++k)
                if (a.s[k] + a.d[k] <= b.s[k] || b.s[k] + b.d[k] <= a.s[k]) return false;
            return true;
        };
        for (size_t i0 = 0; i0 < list.size();) {
            kh::AdjSeedArgs<T> q{};
            int blocks = 0, nfmax = 1;
            size_t i = i0;
            for (; i < list.size() && q.nmon < kh::MAX_DFT; ++i) {
                bool clash = false;
                for (size_t r = i0; r < i && !clash; ++r) clash = meet(*mons[list[r]], *mons[list[i]]);
                if (clash) break;
                const Dft &m = *mons[list[i]];
                kh::AdjSeed<T> &x = q.mon[q.nmon++];
                x.g = adj_cot + off[list[i]];
                x.lam = adjA[electric ? 1 : 0][m.comp % 3];
                x.freqs = m.freqs;
                for (int k = 0; k < 3; ++k) { x.s[k] = m.s[k]; x.d[k] = m.d[k]; }
                x.nf = m.nf;
                x.dec = m.dec;
                x.first = blocks;
                blocks += (int)((m.ncell + 255) / 256);
                nfmax = std::max(nfmax, m.nf);
            }
            q.clk = clk;
            q.koff = koff;
            q.dt = dt;
            q.electric = electric ? 1 : 0;
            q.n0 = g.n[0]; q.n1 = g.n[1]; q.n2 = g.n[2];
            q.sy = g.sy();
            q.sz = g.sz();
            i0 = i;
            if (blocks == 0) continue;
            hipLaunchKernelGGL((kh::adjoint_seed_kernel<T>), dim3(blocks), dim3(256), (size_t)nfmax * 2 * sizeof(double),
                               ast ? ast : st, q);
            HIPCHK(hipGetLastError());
        }
    }

    // The transposed half step of group e (1: E); with `mons`, the group's
    // objective seeds of the step at device clock + koff are added by the
    // same launch (false: too many monitors for one launch, seed separately).
    bool launch_adj_half(int e, const int (*glo)[3], const int (*gn)[3], const size_t *goff, const T *seg,
                         const std::vector<const Dft *> *mons = nullptr, const std::vector<size_t> *off = nullptr,
                         int64_t koff = 0, const double *swrow = nullptr, const std::vector<int> *woff = nullptr) {
        kh::AdjArgs<T> q{};
        if (mons) {
            for (size_t i = 0; i < mons->size(); ++i) {
                const Dft &m = *(*mons)[i];
                if ((m.comp < 3) != (e == 1)) continue;
                if (q.nseed == kh::MAX_DFT) return false;
                kh::AdjSeed<T> &x = q.seed[q.nseed++];
                x.g = adj_cot + (*off)[i];
                x.lam = adjA[e][m.comp % 3];
                x.freqs = m.freqs;
                for (int k = 0; k < 3; ++k) { x.s[k] = m.s[k]; x.d[k] = m.d[k]; }
                x.nf = m.nf;
                x.dec = m.dec;
                x.woff = woff ? (*woff)[i] : 0;
            }
            q.clk = clk;
            q.koff = koff;
            q.sw = swrow;
        }
        for (int a = 0; a < 3; ++a) {
            q.lA[a] = adjA[e][a];
            q.lT[a] = adjT[e][a];
            q.lU[a] = adjU[e][a];
            q.lW[a] = adjW[e][a];
            q.Kout[a] = adjK[e][a];
            q.Kin[a] = adjK[1 - e][a];
            q.sig[a] = e ? sigD[a] : sigB[a];
            q.minv[a] = e ? epsi[a] : mui[a];
            q.inv[a] = T(1) / dx[a];
            q.n[a] = g.n[a];
            if (e) {
                const bool has = gn[a][0] > 0 && gn[a][1] > 0 && gn[a][2] > 0;
                for (int k = 0; k < 3; ++k) { q.glo[a][k] = glo[a][k]; q.gn[a][k] = gn[a][k]; }
                q.fwd[a] = has ? seg + goff[a] : nullptr;
                q.grad[a] = has ? adj_grad + goff[a] : nullptr;
            }
        }
        q.minv_s = e ? eps_s : mu_s;
        q.dt = dt;
        q.sy = g.sy();
        q.sz = g.sz();
        const int64_t ncell = (int64_t)g.n[0] * g.n[1] * g.n[2];
        const unsigned blocks = (unsigned)((ncell + 255) / 256);
        hipStream_t s = ast ? ast : st;
        if (e) hipLaunchKernelGGL((kh::adjoint_half_kernel<T, true>), dim3(blocks), dim3(256), 0, s, q);
        else hipLaunchKernelGGL((kh::adjoint_half_kernel<T, false>), dim3(blocks), dim3(256), 0, s, q);
        HIPCHK(hipGetLastError());
        return true;
    }

    // dJ/d(eps^-1) of a real objective J of DFT monitors, by the discrete
    // transpose of the steps run since timestep 0 (checkpointed), for the
    // three E components over one box each.
    void adjoint_gradient(int nmon, const int32_t *ids, const double *cot, const int32_t *lo, const int32_t *bn,
                          double *grad) override {
        require(ckpt_every > 0, "checkpoints are off: call kh_set_checkpoints before the forward run", KH_E_STATE);
        require(ts > 0, "no steps to differentiate", KH_E_STATE);
        require(nparts == 1, "the adjoint needs real fields (no complex / Bloch fields)");
        require(!mat, "the adjoint does not cover dispersive or chi3 media");
        for (int c = 0; c < 6; ++c) require(!msig[c], "the adjoint does not cover conductive media");
        for (int a = 0; a < 3; ++a) require(!periodic[a], "the adjoint does not cover periodic boundaries");
        require(nmon >= 0 && (nmon == 0 || (ids && cot)), "null monitor list or cotangent");
        const int64_t N = ts, K = ckpt_every;
        const int64_t nseg = (N + K - 1) / K;
        require((int64_t)ck_host.size() >= nseg, "a checkpoint of this run is missing", KH_E_STATE);
        for (int64_t k = 0; k < nseg; ++k)
            require(ck_host[k].valid && ck_host[k].ts == k * K,
                    "a checkpoint of this run is missing (the run did not start at timestep 0 with checkpoints on, "
                    "or state was written during it)",
                    KH_E_STATE);
        std::vector<const Dft *> mons;
        std::vector<size_t> off;
        size_t cot_len = 0;
        for (int i = 0; i < nmon; ++i) {
            require(ids[i] >= 0 && ids[i] < (int)dfts.size(), "monitor id out of range");
            const Dft &m = dfts[ids[i]];
            require(m.comp < 6, "the adjoint takes objectives of E and H monitors (not D or B)");
            mons.push_back(&m);
            off.push_back(cot_len);
            cot_len += (size_t)m.ncell * m.nf * 2;
        }
        // design boxes: inside cells 1..N
        HIPCHK(hipStreamSynchronize(st));
        int glo[3][3], gn[3][3];
        size_t goff[3], gtot = 0;
        for (int a = 0; a < 3; ++a) {
            bool has = true;
            for (int k = 0; k < 3; ++k) {
                glo[a][k] = lo ? lo[3 * a + k] : 1;
                gn[a][k] = bn ? bn[3 * a + k] : 0;
                require(gn[a][k] >= 0, "design box extent must be >= 0");
                has = has && gn[a][k] > 0;
            }
            goff[a] = gtot;
            if (!has) {
                for (int k = 0; k < 3; ++k) gn[a][k] = 0;
                continue;
            }
            for (int k = 0; k < 3; ++k)
                require(glo[a][k] >= 1 && glo[a][k] + gn[a][k] - 1 <= g.n[k], "design box outside cells 1..N");
            gtot += (size_t)gn[a][0] * gn[a][1] * gn[a][2];
        }
        if (!adjA[0][0]) {
            for (int e = 0; e < 2; ++e)
                for (int a = 0; a < 3; ++a) {
                    adjA[e][a] = field_alloc();
                    adjT[e][a] = field_alloc();
                    adjU[e][a] = field_alloc();
                    adjW[e][a] = field_alloc();
                    adjK[e][a] = field_alloc();
                }
        }
        const size_t fb = (size_t)g.total * sizeof(T);
        for (int e = 0; e < 2; ++e)
            for (int a = 0; a < 3; ++a) {
                T *arrs[5] = {adjA[e][a], adjT[e][a], adjU[e][a], adjW[e][a], adjK[e][a]};
                for (T *q : arrs) HIPCHK(hipMemsetAsync(q - g.front, 0, fb, st));
            }
        adj_cot = (double *)regrow(adj_cot, adj_cot_cap, std::max<size_t>(cot_len, 1) * sizeof(double));
        if (cot_len) HIPCHK(hipMemcpyAsync(adj_cot, cot, cot_len * sizeof(double), hipMemcpyHostToDevice, st));
        adj_grad = (double *)regrow(adj_grad, adj_grad_cap, std::max<size_t>(gtot, 1) * sizeof(double));
        HIPCHK(hipMemsetAsync(adj_grad, 0, std::max<size_t>(gtot, 1) * sizeof(double), st));
        // The seeds' weights for the steps of an interval, formed by one
        // launch before its walk back (while each group's seeds go with its
        // transposed half step and the table stays within 64 MiB;
        // KH_SEED_TABLE=0: never): the transposed half steps then read them
        // instead of forming sines and cosines from the device clock per
        // lane.
        std::vector<int> sw_off(mons.size(), 0);
        int sw_stride = 0;
        bool sw_on = !mons.empty();
        {
            int ne = 0, nh = 0;
            for (size_t i = 0; i < mons.size(); ++i) {
                sw_off[i] = sw_stride;
                sw_stride += 1 + 2 * mons[i]->nf;
                ++(mons[i]->comp < 3 ? ne : nh);
            }
            sw_on = sw_on && ne <= kh::MAX_DFT && nh <= kh::MAX_DFT && (int)mons.size() <= kh::MAX_DFT &&
                    (double)K * sw_stride * sizeof(double) <= 64.0 * (1 << 20) && env_int("KH_SEED_TABLE", 1) != 0;
        }
        // Two streams: the walk back of an interval (transposed half steps on
        // st_adj) runs while the next earlier interval is re-run (st), with
        // the snapshots and seed weights in two buffers by interval parity.
        // The walk back reads no device clock (its seeds take the weight
        // table), no forward field and nothing the re-run writes but its own
        // buffer pair; KH_SWEEP_STREAMS=0: one stream.
        const bool pipe = env_int("KH_SWEEP_STREAMS", 1) != 0 && (mons.empty() || sw_on);
        const size_t nbuf = pipe ? 2 : 1;
        adj_seg = (T *)regrow(adj_seg, adj_seg_cap, std::max<size_t>(nbuf * gtot * (size_t)K, 1) * sizeof(T));
        if (sw_on)
            adj_sw = (double *)regrow(adj_sw, adj_sw_cap, nbuf * (size_t)K * sw_stride * sizeof(double));
        if (pipe) {
            if (!st_adj) st_adj = pool_stream(dev);
            for (auto &r : adj_ev)
                for (auto &e : r)
                    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        auto launch_seed_w = [&](int64_t len, int par) {
            kh::DftW<T> w{};
            w.nmon = (int)mons.size();
            for (size_t i = 0; i < mons.size(); ++i) {
                w.freqs[i] = mons[i]->freqs;
                w.nf[i] = mons[i]->nf;
                w.dec[i] = mons[i]->dec;
                w.woff[i] = sw_off[i];
                w.elec[i] = mons[i]->comp < 3 ? 1 : 0;
            }
            w.clk = clk;
            w.koff0 = -len;
            w.len = (int)len;
            w.stride = sw_stride;
            w.dt = dt;
            w.out = adj_sw + (size_t)par * (size_t)K * sw_stride;
            const int64_t n = len * sw_stride;
            hipLaunchKernelGGL((kh::dft_weights_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w);
            HIPCHK(hipGetLastError());
        };
        adj_final = regrow(adj_final, adj_final_cap, state_bytes());
        copy_state(adj_final, nullptr, true);
        const CkHost fin{ts, active, last_active, true};
        HIPCHK(hipStreamSynchronize(st));

        // re-runs skip the monitors; the forward state and monitors are put
        // back however the sweep ends
        std::vector<Dft> saved;
        saved.swap(dfts);
        drop_graph();
        struct Back {
            Ctx *c;
            std::vector<Dft> *saved;
            const CkHost *fin;
            ~Back() {
                try {
                    c->restore_state(c->adj_final, *fin);
                } catch (...) {
                }
                c->end_pre();
                c->dfts.swap(*saved);
                c->plan_graph_len();
                c->drop_graph();
            }
        } back{this, &saved, &fin};
        plan_pre();

        kh::BoxCopy<T> bc{};
        unsigned bmax = 1;
        for (int a = 0; a < 3; ++a) {
            bc.f[a] = fp[0].E[a];
            bc.w[a] = fp[0].WD[a];
            bc.sig[a] = sigD[a];
            for (int k = 0; k < 3; ++k) { bc.lo[a][k] = glo[a][k]; bc.n[a][k] = gn[a][k]; }
            const size_t cnt = (size_t)gn[a][0] * gn[a][1] * gn[a][2];
            bmax = std::max(bmax, (unsigned)((cnt + 255) / 256));
        }
        bc.sy = g.sy();
        bc.sz = g.sz();
        // One interval of len steps from the restored checkpoint: the re-run
        // (snapshot i of the design boxes after step i), then the walk back
        // from the last step; the seeds read the device clock, which the
        // re-run leaves at the interval's end.  Eager: host switch and one
        // tick per step.  Captured: the launches of a whole interval, replayed
        // for every interval of that length (the device clock drives every
        // time-dependent quantity, the snapshots go to the same slots).
        // the re-run's E update writes the snapshots itself where it can
        const bool fused_snap = gtot && snap_fusable(glo, gn);
        struct SnapOff {
            Ctx *c;
            ~SnapOff() {
                for (int a = 0; a < 3; ++a) c->snap_out[a] = nullptr;
            }
        } snap_off{this};
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) {
                snap_lo[a][k] = glo[a][k];
                snap_n[a][k] = gn[a][k];
            }
        auto seg_of = [&](int par, int64_t i) { return adj_seg + ((size_t)par * (size_t)K + (size_t)i) * gtot; };
        // the re-run of one interval of len steps from the restored
        // checkpoint (snapshot i of the design boxes after step i), then the
        // seed weights of its steps (the device clock is at its end)
        auto rerun = [&](int64_t len, bool captured, int par) {
            if (captured) launch_pre((int)len);
            for (int64_t i = 0; i < len; ++i) {
                if (fused_snap)
                    for (int a = 0; a < 3; ++a) snap_out[a] = gn[a][0] > 0 ? seg_of(par, i) + goff[a] : nullptr;
                if (!captured) {
                    begin_step();
                    launch_pre(1);
                }
                launch_step(captured ? (int)i : 0);
                for (int a = 0; a < 3; ++a) snap_out[a] = nullptr;
                if (!captured) {
                    end_pre();
                    launch_tick(1);
                    end_step();
                }
                if (gtot && !fused_snap) {
                    for (int a = 0; a < 3; ++a) bc.out[a] = gn[a][0] > 0 ? seg_of(par, i) + goff[a] : nullptr;
                    hipLaunchKernelGGL((kh::box_copy_kernel<T>), dim3(bmax, 3), dim3(256), 0, st, bc);
                    HIPCHK(hipGetLastError());
                }
            }
            if (captured) {
                end_pre();
                launch_tick(len);
            }
            if (sw_on) launch_seed_w(len, par);
        };
        // the walk back from the interval's last step, on ast
        auto walk = [&](int64_t len, int par) {
            const double *swb = sw_on ? adj_sw + (size_t)par * (size_t)K * sw_stride : nullptr;
            for (int64_t i = len - 1; i >= 0; --i) {
                // the seeds go with the transposed half step (or before it,
                // as their own launches, past MAX_DFT monitors)
                const double *swrow = sw_on ? swb + (size_t)i * sw_stride : nullptr;
                if (!launch_adj_half(1, glo, gn, goff, seg_of(par, i), &mons, &off, i - len, swrow, &sw_off)) {
                    launch_adj_seed(mons, off, true, i - len);
                    launch_adj_half(1, glo, gn, goff, seg_of(par, i));
                }
                if (!launch_adj_half(0, glo, gn, goff, nullptr, &mons, &off, i - len, swrow, &sw_off)) {
                    launch_adj_seed(mons, off, false, i - len);
                    launch_adj_half(0, glo, gn, goff, nullptr);
                }
            }
        };
        struct AstOff {
            Ctx *c;
            ~AstOff() { c->ast = nullptr; }
        } ast_off{this};
        ast = pipe ? st_adj : st;
        // graphs of a full interval: one stream, re-run and walk back
        // together, with and without the FAST source correction (captured
        // with it while the sources are or were on at the interval's start;
        // the kernels gate themselves by the clock); two streams, the re-run
        // per source state and buffer, the walk back per buffer.  Destroyed
        // after both streams have drained, however the sweep ends.
        struct SweepGraphs {
            hipStream_t st, st2;
            hipGraph_t g[6]{};
            hipGraphExec_t x[6]{};
            ~SweepGraphs() {
                (void)hipStreamSynchronize(st);
                if (st2) (void)hipStreamSynchronize(st2);
                for (int i = 0; i < 6; ++i) {
                    if (x[i]) (void)hipGraphExecDestroy(x[i]);
                    if (g[i]) (void)hipGraphDestroy(g[i]);
                }
            }
        } sw{st, pipe ? st_adj : nullptr};
        bool use_graph = graph_steps > 0 && !timing && !graph_failed && K <= 1024;
        // slot: one stream 0/1 (source state); two streams 0..3 re-run
        // (source state * 2 + buffer), 4/5 walk back (buffer)
        auto capture = [&](int slot, hipStream_t cs, const std::function<void()> &body) -> bool {
            HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            bool ok = true;
            try {
                body();
            } catch (const KhError &) {
                ok = false;
            }
            end_pre();
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(cs, &gr);
            if (!ok || e != hipSuccess || !gr) {
                if (gr) (void)hipGraphDestroy(gr);
                (void)hipGetLastError();
                return false;
            }
            sw.g[slot] = gr;
            if (hipGraphInstantiate(&sw.x[slot], gr, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGetLastError();
                sw.x[slot] = nullptr;
                return false;
            }
            return true;
        };
        bool used[2] = {false, false};
        for (int64_t sgi = nseg - 1; sgi >= 0; --sgi) {
            const int64_t s0 = sgi * K, s1 = std::min(N, s0 + K), len = s1 - s0;
            const int par = pipe ? (int)(sgi & 1) : 0;
            restore_state(ckpt[(size_t)sgi], ck_host[(size_t)sgi]);
            // the walk back that read this buffer pair has finished
            if (pipe && used[par]) HIPCHK(hipStreamWaitEvent(st, adj_ev[1][par], 0));
            const int which = (active || last_active) ? 1 : 0;
            bool ran = false;
            if (use_graph && len == K) {
                const int slot = pipe ? which * 2 + par : which;
                if (!sw.x[slot] && !capture(slot, st, [&] {
                        if (pipe) rerun(K, true, par);
                        else {
                            rerun(K, true, 0);
                            walk(K, 0);
                        }
                    }))
                    use_graph = false;
                if (use_graph) {
                    HIPCHK(hipGraphLaunch(sw.x[slot], st));
                    for (int64_t j = s0; j < s1; ++j) {
                        begin_step();
                        end_step();
                    }
                    ran = true;
                }
            }
            if (!ran) {
                rerun(len, false, par);
                if (!pipe) walk(len, 0);
            }
            if (!pipe) continue;
            HIPCHK(hipEventRecord(adj_ev[0][par], st));
            HIPCHK(hipStreamWaitEvent(st_adj, adj_ev[0][par], 0));
            bool walked = false;
            if (use_graph && len == K) {
                const int slot = 4 + par;
                if (!sw.x[slot] && !capture(slot, st_adj, [&] { walk(K, par); })) use_graph = false;
                if (use_graph) {
                    HIPCHK(hipGraphLaunch(sw.x[slot], st_adj));
                    walked = true;
                }
            }
            if (!walked) walk(len, par);
            HIPCHK(hipEventRecord(adj_ev[1][par], st_adj));
            used[par] = true;
        }
        if (pipe) HIPCHK(hipStreamSynchronize(st_adj));
        if (gtot) HIPCHK(hipMemcpyAsync(grad, adj_grad, gtot * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }

    void reset() override {
        HIPCHK(hipStreamSynchronize(st));
        size_t fb = (size_t)g.total * sizeof(T);
        for (int part = 0; part < nparts; ++part)
            for (int a = 0; a < 3; ++a) {
                const Part &f = fp[part];
                T *arrs[10] = {f.E[a], f.H[a], f.D[a], f.B[a], f.UD[a], f.UB[a], f.WD[a], f.WB[a], f.CD[a], f.CB[a]};
                for (T *p : arrs)
                    if (p) HIPCHK(hipMemsetAsync(p - g.front, 0, fb, st));
            }
        for (auto &m : dfts)
            HIPCHK(hipMemsetAsync(m.acc, 0, (size_t)m.ncell * m.nf * 2 * (acc64 ? sizeof(double) : sizeof(T)), st));
        for (auto &q : poles) {
            const size_t nb = (size_t)q.n[0] * q.n[1] * q.n[2] * sizeof(T);
            for (int part = 0; part < nparts; ++part)
                for (int a = 0; a < 3; ++a)
                    if (q.P[part][a]) {
                        HIPCHK(hipMemsetAsync(q.P[part][a], 0, nb, st));
                        HIPCHK(hipMemsetAsync(q.Pp[part][a], 0, nb, st));
                    }
        }
        clear_material_state();
        HIPCHK(hipMemsetAsync(clk, 0, 256, st));
        HIPCHK(hipStreamSynchronize(st));
        for (auto &c : ck_host) c.valid = false;
        ts = 0;
        active = true;
        last_active = false;
    }

    void clear_material_state() {
        const size_t nb = (size_t)mcells * sizeof(T);
        for (int part = 0; part < 2; ++part)
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 2; ++b)
                    if (fpd[b][part][a]) HIPCHK(hipMemsetAsync(fpd[b][part][a], 0, nb, st));
                if (delta[part][a]) HIPCHK(hipMemsetAsync(delta[part][a], 0, nb, st));
            }
    }

    void set_timing(int n) override {
        timing = n != 0;
        while ((int64_t)pool.size() < (int64_t)n) {
            EvPair e{};
            HIPCHK(hipEventCreate(&e.a));
            HIPCHK(hipEventCreate(&e.b));
            pool.push_back(e);
        }
    }

    void get_timing(double *out) override {
        HIPCHK(hipStreamSynchronize(st));
        for (auto &e : evs) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
            tsum[e.kind] += ms;
            tcnt[e.kind] += 1;
            pool.push_back(e);
        }
        evs.clear();
        for (int k = 0; k < kh_timing_slots; ++k) {
            out[2 * k] = tsum[k];
            out[2 * k + 1] = tcnt[k];
            tsum[k] = 0;
            tcnt[k] = 0;
        }
    }

    void *stream() override { return (void *)st; }
};

size_t footprint(const kh_sim_desc &d) {
    Grid g;
    g.init(d.nx, d.ny, d.nz);
    size_t es = d.precision == KH_F32 ? 4 : 8;
    // E, H, U and W per component (per part with complex fields); D and B only in EXACT mode
    size_t arrays = (d.mode == KH_MODE_EXACT ? 24 : 18) * (d.complex_fields ? 2 : 1);
    for (int a = 0; a < 3; ++a) arrays += (d.eps_inv[a] != nullptr) + (d.mu_inv[a] != nullptr);
    // byte-per-cell material indices (when the values fit a table)
    size_t idx_bytes = 0;
    for (int a = 0; a < 3; ++a) idx_bytes += ((d.eps_inv[a] != nullptr) + (d.mu_inv[a] != nullptr)) * (size_t)g.total;
    // conductivity: sigma plus C per part
    for (int c = 0; c < 6; ++c) arrays += (d.sigma_mat[c] != nullptr) * (1 + (d.complex_fields ? 2 : 1));
    size_t sig = 0;
    for (int a = 0; a < 3; ++a) sig += 2 * ((size_t)g.n[a] + 2) * es;
    return arrays * (size_t)g.total * es + idx_bytes + sig + 256 * 64 * sizeof(double);
}

void validate(const kh_sim_desc *d) {
    require(d != nullptr, "desc is null");
    require(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "grid extents must be >= 1");
    require(d->precision == KH_F32 || d->precision == KH_F64, "precision must be KH_F32 or KH_F64");
    require(d->mode == KH_MODE_FAST || d->mode == KH_MODE_EXACT, "mode must be KH_MODE_FAST or KH_MODE_EXACT");
    require(d->dt > 0 && d->dx > 0 && d->dy > 0 && d->dz > 0, "dt and grid spacings must be > 0");
    int nb = 0;
    for (int a = 0; a < 3; ++a) nb += (d->sigma_b[a] != nullptr) + (d->sigma_d[a] != nullptr);
    // 2D (nz = 1, dz = inf) has no z profiles (src/Boundaries.jl:130-162)
    const bool twod = d->nz == 1 && std::isinf(d->dz) && d->sigma_b[2] == nullptr && d->sigma_d[2] == nullptr;
    require(nb == 0 || nb == 6 || (twod && nb == 4),
            "sigma_b/sigma_d: give all six profiles, none, or (2D: nz = 1, dz = inf) the four x/y ones");
    if (nb == 4) nb = 6;
    int nc = 0;
    for (int c = 0; c < 6; ++c) nc += d->sigma_mat[c] != nullptr;
    // the reference has no curl update for a conductive component without PML
    // arrays (src/Kernels/Helpers.jl:141-151 and :225-236 are ambiguous there)
    require(nc == 0 || nb == 6,
            "material conductivity needs the six PML profiles (zeros where there is no PML)");
}

}  // namespace

struct kh_ctx {
    std::unique_ptr<CtxBase> impl;
};

// Continuous-batching scheduler over contexts (the serving form of
// src/Batch.jl:142-263, admission by footprint as plan_batch, :22-69).
// admit/submit may come from preparation threads while one thread polls;
// every context is stepped from the polling thread only.
struct kh_server {
    // chunks queued ahead on each context's stream: the next chunk waits on
    // the stream while the current one runs, so a context does not drain
    // between two polls
    static constexpr int DEPTH = 2;
    struct Req {
        kh_ctx *ctx;
        int64_t remaining;
        double footprint;
        int64_t id;
        std::deque<hipEvent_t> pending;  // one event after each queued chunk
    };
    double budget = 0.0, used = 0.0;
    int max_active = 1, chunk = 256;
    int reserved = 0;           // admitted, context not yet submitted
    std::deque<Req> active;     // submitted, stepping or waiting for a chunk
    std::deque<int64_t> done;   // finished, not yet reported
    std::vector<hipEvent_t> events;  // idle events for reuse
    std::mutex mu;
    hipEvent_t take_event() {
        if (!events.empty()) {
            hipEvent_t e = events.back();
            events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }
    ~kh_server() {
        for (auto &r : active)
            for (hipEvent_t e : r.pending) (void)hipEventDestroy(e);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

namespace {
// Host copy of an object list in the kernels' form; prism vertex lists are
// gathered into one array (one device copy).
struct RasterInput {
    std::vector<kh::RasterObj> ho;
    std::vector<double> verts;
    std::vector<size_t> vfirst;
};

RasterInput pack_objects(const kh_raster_object *objs, int n_objs) {
    RasterInput in;
    in.ho.resize(std::max(1, n_objs));
    in.vfirst.assign(std::max(1, n_objs), 0);
    for (int o = 0; o < n_objs; ++o) {
        const kh_raster_object &r = objs[o];
        require(r.shape >= 0 && r.shape <= 3, "shape must be KH_SHAPE_BOX, _SPHERE, _CYLINDER or _PRISM");
        require(r.axis >= 0 && r.axis <= 2, "cylinder axis must be 0..2");
        in.ho[o] = kh::RasterObj{};
        if (r.shape == KH_SHAPE_PRISM) {
            require(r.n_vertices >= 3 && r.n_vertices <= 1000000 && r.vertices != nullptr,
                    "a prism needs at least 3 vertices");
            in.vfirst[o] = in.verts.size();
            in.verts.insert(in.verts.end(), r.vertices, r.vertices + 2 * (size_t)r.n_vertices);
            in.ho[o].nv = r.n_vertices;
        }
        in.ho[o].shape = r.shape;
        in.ho[o].axis = r.axis;
        for (int a = 0; a < 3; ++a) {
            in.ho[o].c[a] = r.center[a];
            in.ho[o].p[a] = r.size[a];
        }
        for (int c = 0; c < 6; ++c) in.ho[o].value[c] = r.value[c];
    }
    return in;
}

constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// Device staging of one call: objects | coordinates | prism vertices |
// `rasters` component arrays of `ab` bytes, each part 256-byte aligned.
struct RasterStage {
    size_t ob, cb, vb, ab, total;
    RasterStage(const RasterInput &in, const int32_t *cnt, size_t elem, int rasters) {
        ob = in.ho.size() * sizeof(kh::RasterObj);
        cb = (size_t)(cnt[0] + cnt[1] + cnt[2]) * sizeof(double);
        vb = in.verts.size() * sizeof(double);
        ab = (size_t)cnt[0] * cnt[1] * cnt[2] * elem;
        total = round256(ob) + round256(cb) + round256(vb) + rasters * round256(ab);
    }
};

// Copies the inputs into `stage` (device memory of RasterStage::total bytes)
// on stream `st` and paints the component into `paint`; with smoothing the
// smoothed values go to `smoothed` (then `paint` is the unsmoothed raster).
template <typename T>
void raster_launch(RasterInput &in, const RasterStage &L, char *stage, int n_objs, int comp, const int32_t *cnt,
                   const double *xs, const double *ys, const double *zs, bool smooth, int smoothing,
                   const double *half, T *paint, T *smoothed, hipStream_t st) {
    kh::RasterObj *d_obj = (kh::RasterObj *)stage;
    double *d_x = (double *)(stage + round256(L.ob));
    double *d_y = d_x + cnt[0], *d_z = d_y + cnt[1];
    double *d_v = (double *)((char *)d_x + round256(L.cb));
    for (int o = 0; o < n_objs; ++o)
        if (in.ho[o].shape == KH_SHAPE_PRISM) in.ho[o].v = d_v + in.vfirst[o];
    if (L.vb) HIPCHK(hipMemcpyAsync(d_v, in.verts.data(), L.vb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_obj, in.ho.data(), L.ob, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_x, xs, cnt[0] * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_y, ys, cnt[1] * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_z, zs, cnt[2] * sizeof(double), hipMemcpyHostToDevice, st));
    const int64_t n = (int64_t)cnt[0] * cnt[1] * cnt[2];
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(kh::raster_kernel<T>, dim3(blocks), dim3(256), 0, st, (const kh::RasterObj *)d_obj, n_objs,
                       comp, cnt[0], cnt[1], cnt[2], (const double *)d_x, (const double *)d_y, (const double *)d_z,
                       paint);
    HIPCHK(hipGetLastError());
    if (smooth) {
        hipLaunchKernelGGL(kh::smooth_kernel<T>, dim3(blocks), dim3(256), 0, st, (const kh::RasterObj *)d_obj,
                           n_objs, comp, smoothing == 2 ? 1 : 0, cnt[0], cnt[1], cnt[2], (const double *)d_x,
                           (const double *)d_y, (const double *)d_z, half[0], half[1], half[2], (const T *)paint,
                           smoothed);
        HIPCHK(hipGetLastError());
    }
}

template <typename T>
void raster_run(const kh_raster_object *objs, int n_objs, int comp, const int32_t *cnt, const double *xs,
                const double *ys, const double *zs, int smoothing, const double *half, void *out) {
    RasterInput in = pack_objects(objs, n_objs);
    const bool smooth = smoothing != 0 && comp < 3;
    const RasterStage L(in, cnt, sizeof(T), smooth ? 2 : 1);
    void *buf = nullptr;
    HIPCHK(hipMalloc(&buf, L.total));
    struct Free {
        void *p;
        ~Free() { (void)hipFree(p); }
    } guard{buf};
    char *stage = (char *)buf;
    T *d_a = (T *)(stage + L.total - (smooth ? 2 : 1) * round256(L.ab));
    T *d_s = smooth ? (T *)((char *)d_a + round256(L.ab)) : nullptr;
    raster_launch<T>(in, L, stage, n_objs, comp, cnt, xs, ys, zs, smooth, smoothing, half, d_a, d_s, 0);
    HIPCHK(hipMemcpy(out, smooth ? d_s : d_a, L.ab, hipMemcpyDeviceToHost));
}

}  // namespace

// The allocation-free variant's staging: a device array of fixed size; every
// GPU holds its own copy (at its own address) until the process ends.  Host
// code never passes `scratch` itself (a host-side stand-in), only the device
// address looked up once per device below.
static __device__ __align__(256) unsigned char scratch[KH_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_dev[256] = {};

void *scratch_on_current_device() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    require(dev >= 0 && dev < 256, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw KhError(KH_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

template <typename T>
void raster_run_noalloc(const kh_raster_object *objs, int n_objs, int comp, const int32_t *cnt, const double *xs,
                        const double *ys, const double *zs, int smoothing, const double *half, void *out_dev,
                        hipStream_t st) {
    RasterInput in = pack_objects(objs, n_objs);
    const bool smooth = smoothing != 0 && comp < 3;
    // staging: the inputs and, with smoothing, the unsmoothed raster; the
    // painted (or smoothed) values go straight to the caller's array
    const RasterStage L(in, cnt, sizeof(T), smooth ? 1 : 0);
    require(L.total <= (size_t)KH_NOALLOC_SCRATCH_BYTES,
            "kh_raster_component_noalloc: the call needs " + std::to_string(L.total) +
                " bytes of staging, more than the " + std::to_string((size_t)KH_NOALLOC_SCRATCH_BYTES) +
                " bytes of the library's scratch array; use kh_raster_component");
    char *stage = (char *)scratch_on_current_device();
    T *paint = smooth ? (T *)(stage + L.total - round256(L.ab)) : (T *)out_dev;
    raster_launch<T>(in, L, stage, n_objs, comp, cnt, xs, ys, zs, smooth, smoothing, half, paint,
                     smooth ? (T *)out_dev : nullptr, st);
    // the host inputs (object list, coordinates) are read by the copies above
    HIPCHK(hipStreamSynchronize(st));
}
}  // namespace

extern "C" {

const char *kh_last_error(void) { return g_last_error.c_str(); }

int kh_version(void) { return KH_ABI_VERSION; }

int kh_mem_footprint(const kh_sim_desc *desc, size_t *bytes) {
    return guarded([&] {
        validate(desc);
        require(bytes != nullptr, "bytes is null");
        *bytes = footprint(*desc);
    });
}

int kh_create(const kh_sim_desc *desc, kh_ctx **out) {
    return guarded([&] {
        validate(desc);
        require(out != nullptr, "out is null");
        *out = nullptr;
        auto c = std::make_unique<kh_ctx>();
        if (desc->precision == KH_F32) c->impl = std::make_unique<Ctx<float>>(*desc);
        else c->impl = std::make_unique<Ctx<double>>(*desc);
        *out = c.release();
    });
}

int kh_destroy(kh_ctx *ctx) {
    return guarded([&] { delete ctx; });
}

#define CTX_OK require(ctx != nullptr && ctx->impl != nullptr, "ctx is null")

int kh_add_source(kh_ctx *ctx, int32_t component, const int32_t start_idx[3], const int32_t dims[3],
                  const void *amplitude, int32_t *source_id) {
    return guarded([&] {
        CTX_OK;
        require(start_idx && dims && amplitude && source_id, "null argument");
        ctx->impl->add_source(component, start_idx, dims, amplitude, source_id);
    });
}

int kh_set_waveform(kh_ctx *ctx, int32_t source_id, int64_t n_steps, const double *values) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->set_waveform(source_id, n_steps, values);
    });
}

int kh_add_dft(kh_ctx *ctx, int32_t component, const int32_t start_idx[3], const int32_t dims[3],
               int32_t n_freq, const double *freqs, int32_t decimation, int32_t *dft_id) {
    return guarded([&] {
        CTX_OK;
        require(start_idx && dims && freqs && dft_id, "null argument");
        ctx->impl->add_dft(component, start_idx, dims, n_freq, freqs, decimation, dft_id);
    });
}

int kh_set_eps_inv(kh_ctx *ctx, int32_t component, const void *eps_inv) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->set_eps_inv(component, eps_inv);
    });
}

int kh_set_chi3(kh_ctx *ctx, const void *chi3) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->set_chi3(chi3);
    });
}

int kh_add_pole(kh_ctx *ctx, const double coeffs[5], int32_t is_drude, const void *sigma, int32_t *pole_id) {
    return guarded([&] {
        CTX_OK;
        require(pole_id != nullptr, "pole_id is null");
        ctx->impl->add_pole(coeffs, is_drude, sigma, pole_id);
    });
}

int kh_read_pole(kh_ctx *ctx, int32_t pole_id, int32_t which, int32_t component, void *host, size_t bytes) {
    return guarded([&] {
        CTX_OK;
        require(host != nullptr, "host buffer is null");
        ctx->impl->read_pole(pole_id, which, component, host, bytes);
    });
}

int kh_write_pole(kh_ctx *ctx, int32_t pole_id, int32_t which, int32_t component, const void *host,
                  size_t bytes) {
    return guarded([&] {
        CTX_OK;
        require(host != nullptr, "host buffer is null");
        ctx->impl->write_pole(pole_id, which, component, host, bytes);
    });
}

int kh_step(kh_ctx *ctx, int64_t n_steps) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->step(n_steps);
    });
}

int kh_set_graph(kh_ctx *ctx, int32_t steps_per_graph) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->set_graph(steps_per_graph);
    });
}

int kh_prepare_graphs(kh_ctx *ctx, int64_t n_steps) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->prepare_graphs(n_steps);
    });
}

int kh_get_launch_counts(kh_ctx *ctx, int64_t counts[4]) {
    return guarded([&] {
        CTX_OK;
        require(counts != nullptr, "null argument");
        ctx->impl->launch_counts(counts);
    });
}

int kh_server_create(double budget_bytes, int32_t max_active, int32_t chunk_steps, kh_server **out) {
    return guarded([&] {
        require(out != nullptr, "out is null");
        require(budget_bytes > 0 && std::isfinite(budget_bytes), "budget must be positive");
        require(max_active >= 1 && chunk_steps >= 1, "max_active and chunk_steps must be >= 1");
        std::unique_ptr<kh_server> s(new kh_server());
        s->budget = budget_bytes;
        s->max_active = max_active;
        s->chunk = chunk_steps;
        *out = s.release();
    });
}

int kh_server_destroy(kh_server *s) {
    return guarded([&] {
        require(s != nullptr, "server is null");
        delete s;
    });
}

int kh_server_admit(kh_server *s, double footprint, int32_t *admitted) {
    return guarded([&] {
        require(s != nullptr && admitted != nullptr, "null argument");
        require(footprint >= 0 && std::isfinite(footprint), "footprint must be >= 0");
        std::lock_guard<std::mutex> lock(s->mu);
        const int inflight = s->reserved + (int)s->active.size();
        // a request larger than the budget still runs when it would run alone
        const bool fits = s->used + footprint <= s->budget || inflight == 0;
        *admitted = (inflight < s->max_active && fits) ? 1 : 0;
        if (*admitted) {
            s->reserved += 1;
            s->used += footprint;
        }
    });
}

int kh_server_submit(kh_server *s, kh_ctx *ctx, int64_t n_steps, double footprint, int64_t request_id) {
    return guarded([&] {
        require(s != nullptr, "server is null");
        CTX_OK;
        require(n_steps >= 0, "n_steps must be >= 0");
        std::lock_guard<std::mutex> lock(s->mu);
        require(s->reserved > 0, "submit without a matching kh_server_admit");
        s->reserved -= 1;
        s->active.push_back({ctx, n_steps, footprint, request_id});
    });
}

int kh_server_cancel(kh_server *s, double footprint) {
    return guarded([&] {
        require(s != nullptr, "server is null");
        std::lock_guard<std::mutex> lock(s->mu);
        require(s->reserved > 0, "cancel without a matching kh_server_admit");
        s->reserved -= 1;
        s->used -= footprint;
    });
}

int kh_server_poll(kh_server *s, int64_t *done_ids, int32_t cap, int32_t *n_done, int32_t *n_inflight) {
    return guarded([&] {
        require(s != nullptr && n_done != nullptr, "null argument");
        require(cap >= 0 && (cap == 0 || done_ids != nullptr), "bad done buffer");
        std::lock_guard<std::mutex> lock(s->mu);
        for (auto it = s->active.begin(); it != s->active.end();) {
            // completed chunks
            while (!it->pending.empty()) {
                const hipError_t e = hipEventQuery(it->pending.front());
                if (e == hipErrorNotReady) break;
                HIPCHK(e);
                s->events.push_back(it->pending.front());
                it->pending.pop_front();
            }
            // keep DEPTH chunks queued (kh_step is asynchronous on the context's stream)
            while (it->remaining > 0 && (int)it->pending.size() < kh_server::DEPTH) {
                // whole main-graph blocks, so that every chunk starts on the
                // same phase of the DFT period and replays the same graph
                const int64_t blk = it->ctx->impl->graph_block();
                const int64_t n = std::min<int64_t>((s->chunk + blk - 1) / blk * blk, it->remaining);
                it->ctx->impl->step(n);
                it->remaining -= n;
                hipEvent_t ev = s->take_event();
                HIPCHK(hipEventRecord(ev, (hipStream_t)it->ctx->impl->stream()));
                it->pending.push_back(ev);
            }
            if (it->remaining > 0 || !it->pending.empty() || !it->ctx->impl->idle()) {
                ++it;
                continue;
            }
            s->used -= it->footprint;
            s->done.push_back(it->id);
            it = s->active.erase(it);
        }
        int k = 0;
        while (k < cap && !s->done.empty()) {
            done_ids[k++] = s->done.front();
            s->done.pop_front();
        }
        *n_done = k;
        if (n_inflight) *n_inflight = s->reserved + (int)s->active.size();
    });
}

int kh_query(kh_ctx *ctx, int32_t *idle) {
    return guarded([&] {
        CTX_OK;
        require(idle != nullptr, "idle is null");
        *idle = ctx->impl->idle() ? 1 : 0;
    });
}

int kh_sync(kh_ctx *ctx) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->sync();
    });
}

int kh_get_timestep(kh_ctx *ctx, int64_t *timestep) {
    return guarded([&] {
        CTX_OK;
        require(timestep != nullptr, "null argument");
        *timestep = ctx->impl->timestep();
    });
}

int kh_sources_active(kh_ctx *ctx, int32_t *active) {
    return guarded([&] {
        CTX_OK;
        require(active != nullptr, "null argument");
        *active = ctx->impl->sources_active() ? 1 : 0;
    });
}

int kh_read_field(kh_ctx *ctx, int32_t component, void *host, size_t bytes) {
    return guarded([&] {
        CTX_OK;
        require(host != nullptr, "null argument");
        ctx->impl->read_field(component, host, bytes);
    });
}

int kh_write_field(kh_ctx *ctx, int32_t component, const void *host, size_t bytes) {
    return guarded([&] {
        CTX_OK;
        require(host != nullptr, "null argument");
        ctx->impl->write_field(component, host, bytes);
    });
}

int kh_read_dft(kh_ctx *ctx, int32_t dft_id, double *host, size_t bytes) {
    return guarded([&] {
        CTX_OK;
        require(host != nullptr, "null argument");
        ctx->impl->read_dft(dft_id, host, bytes);
    });
}

int kh_flux(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis, double dA,
            double *out) {
    return guarded([&] {
        CTX_OK;
        require(out != nullptr, "null argument");
        ctx->impl->flux(e1, e2, h1, h2, normal_axis, dA, out);
    });
}

int kh_mode_overlap(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis,
                    int32_t n1, int32_t n2, const double *mode, double dA, double *out) {
    return guarded([&] {
        CTX_OK;
        require(mode != nullptr && out != nullptr, "null argument");
        ctx->impl->mode_overlap(e1, e2, h1, h2, normal_axis, n1, n2, mode, dA, out);
    });
}

int kh_near2far(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis,
                double normal_sign, const double *base, double d1, double d2, double dA, int32_t n1,
                int32_t n2, const double *freqs, int32_t n_freq, double medium_eps, double medium_mu,
                const double *obs, int32_t n_obs, double *out) {
    return guarded([&] {
        CTX_OK;
        require(base != nullptr && freqs != nullptr && obs != nullptr && out != nullptr, "null argument");
        const int32_t ids[4] = {e1, e2, h1, h2};
        ctx->impl->near2far(ids, normal_axis, normal_sign, base, d1, d2, dA, n1, n2, freqs, n_freq, medium_eps,
                            medium_mu, obs, n_obs, out);
    });
}

int kh_near2far_spectrum(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis,
                         int32_t n1, int32_t n2, int32_t freq_index, int32_t pad1, int32_t pad2,
                         const int32_t *bins, int32_t n_bins, double *out) {
    return guarded([&] {
        CTX_OK;
        require(bins != nullptr && out != nullptr, "null argument");
        const int32_t ids[4] = {e1, e2, h1, h2};
        ctx->impl->near2far_spectrum(ids, normal_axis, n1, n2, freq_index, pad1, pad2, bins, n_bins, out);
    });
}


int kh_raster_component(int32_t precision, int32_t component, const int32_t cnt[3], const double *xs,
                        const double *ys, const double *zs, const kh_raster_object *objs, int32_t n_objs,
                        int32_t smoothing, const double half[3], void *out) {
    return guarded([&] {
        require(precision == KH_F32 || precision == KH_F64, "precision must be KH_F32 or KH_F64");
        require(component >= 0 && component < 6, "component must be 0..5");
        require(cnt != nullptr && xs != nullptr && ys != nullptr && zs != nullptr && out != nullptr, "null argument");
        require(cnt[0] >= 1 && cnt[1] >= 1 && cnt[2] >= 1, "empty component array");
        require((int64_t)cnt[0] * cnt[1] * cnt[2] < (1LL << 40), "component array too large");
        require(n_objs >= 0 && (n_objs == 0 || objs != nullptr), "bad object list");
        require(smoothing >= 0 && smoothing <= 2, "smoothing must be 0 (none), 1 (volume average), 2 (anisotropic)");
        require(smoothing == 0 || component >= 3 || half != nullptr, "smoothing needs the half cell size");
        if (precision == KH_F32)
            raster_run<float>(objs, n_objs, component, cnt, xs, ys, zs, smoothing, half, out);
        else
            raster_run<double>(objs, n_objs, component, cnt, xs, ys, zs, smoothing, half, out);
    });
}

int kh_raster_component_noalloc(int32_t precision, int32_t component, const int32_t cnt[3], const double *xs,
                                const double *ys, const double *zs, const kh_raster_object *objs, int32_t n_objs,
                                int32_t smoothing, const double half[3], void *out_device, void *stream) {
    return guarded([&] {
        require(precision == KH_F32 || precision == KH_F64, "precision must be KH_F32 or KH_F64");
        require(component >= 0 && component < 6, "component must be 0..5");
        require(cnt != nullptr && xs != nullptr && ys != nullptr && zs != nullptr && out_device != nullptr,
                "null argument");
        require(cnt[0] >= 1 && cnt[1] >= 1 && cnt[2] >= 1, "empty component array");
        require((int64_t)cnt[0] * cnt[1] * cnt[2] < (1LL << 40), "component array too large");
        require(n_objs >= 0 && (n_objs == 0 || objs != nullptr), "bad object list");
        require(smoothing >= 0 && smoothing <= 2, "smoothing must be 0 (none), 1 (volume average), 2 (anisotropic)");
        require(smoothing == 0 || component >= 3 || half != nullptr, "smoothing needs the half cell size");
        // out_device must be device memory of the current device holding the
        // whole component array (the kernels write every value)
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipPointerAttribute_t attr{};
        const bool is_dev = hipPointerGetAttributes(&attr, out_device) == hipSuccess &&
                            attr.type == hipMemoryTypeDevice && attr.device == dev;
        (void)hipGetLastError();
        require(is_dev, "out_device must be device memory of the current device");
        hipDeviceptr_t base = nullptr;
        size_t range = 0;
        const size_t need = (size_t)cnt[0] * cnt[1] * cnt[2] * (precision == KH_F32 ? 4 : 8);
        const bool fits = hipMemGetAddressRange(&base, &range, (hipDeviceptr_t)out_device) == hipSuccess &&
                          (const char *)out_device + need <= (const char *)base + range;
        (void)hipGetLastError();
        require(fits, "out_device holds fewer than cnt[0] * cnt[1] * cnt[2] values");
        if (precision == KH_F32)
            raster_run_noalloc<float>(objs, n_objs, component, cnt, xs, ys, zs, smoothing, half, out_device,
                                      (hipStream_t)stream);
        else
            raster_run_noalloc<double>(objs, n_objs, component, cnt, xs, ys, zs, smoothing, half, out_device,
                                       (hipStream_t)stream);
    });
}

int kh_dft_norm(kh_ctx *ctx, int32_t dft_id, double *out) {
    return guarded([&] {
        CTX_OK;
        require(out != nullptr, "null argument");
        ctx->impl->dft_norm(dft_id, out);
    });
}

int kh_dft_norms(kh_ctx *ctx, int32_t n_out, double *out) {
    return guarded([&] {
        CTX_OK;
        require(out != nullptr || n_out == 0, "null argument");
        ctx->impl->dft_norms(n_out, out);
    });
}

int kh_reset(kh_ctx *ctx) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->reset();
    });
}

int kh_set_checkpoints(kh_ctx *ctx, int32_t every) {
    return guarded([&] {
        CTX_OK;
        ctx->impl->set_checkpoints(every);
    });
}

int kh_adjoint_gradient(kh_ctx *ctx, int32_t n_monitors, const int32_t *monitor_ids, const double *cotangent,
                        const int32_t box_lo[9], const int32_t box_n[9], double *grad) {
    return guarded([&] {
        CTX_OK;
        require(box_lo != nullptr && box_n != nullptr, "null design box");
        ctx->impl->adjoint_gradient(n_monitors, monitor_ids, cotangent, box_lo, box_n, grad);
    });
}

int kh_stream_rate(int32_t n_read, int32_t n_write, int64_t elements, int32_t iterations, double *gb_per_s) {
    return guarded([&] {
        require(n_read >= 1 && n_read <= 6, "n_read must be 1..6");
        require(n_write >= 1 && n_write <= 3, "n_write must be 1..3");
        require(elements >= 512 && elements % 2 == 0, "elements must be even and >= 512");
        require(elements / 2 <= (int64_t)256 * ((1LL << 31) - 1), "elements too large");
        require(iterations >= 1 && iterations <= 1000, "iterations must be 1..1000");
        require(gb_per_s != nullptr, "null result");
        using SF = void (*)(kh::StreamArgs);
        static const SF table[6][3] = {
            {kh::stream_kernel<1, 1>, kh::stream_kernel<1, 2>, kh::stream_kernel<1, 3>},
            {kh::stream_kernel<2, 1>, kh::stream_kernel<2, 2>, kh::stream_kernel<2, 3>},
            {kh::stream_kernel<3, 1>, kh::stream_kernel<3, 2>, kh::stream_kernel<3, 3>},
            {kh::stream_kernel<4, 1>, kh::stream_kernel<4, 2>, kh::stream_kernel<4, 3>},
            {kh::stream_kernel<5, 1>, kh::stream_kernel<5, 2>, kh::stream_kernel<5, 3>},
            {kh::stream_kernel<6, 1>, kh::stream_kernel<6, 2>, kh::stream_kernel<6, 3>}};
        const size_t bytes = (size_t)elements * sizeof(float);
        struct Res {
            std::vector<void *> bufs;
            hipStream_t s = nullptr;
            hipEvent_t a = nullptr, b = nullptr;
            ~Res() {
                if (s) (void)hipStreamSynchronize(s);
                for (void *q : bufs) (void)hipFree(q);
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
                if (s) (void)hipStreamDestroy(s);
            }
        } r;
        HIPCHK(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&r.a));
        HIPCHK(hipEventCreate(&r.b));
        kh::StreamArgs p{};
        for (int k = 0; k < n_read + n_write; ++k) {
            void *q = nullptr;
            HIPCHK(hipMalloc(&q, bytes));
            r.bufs.push_back(q);
            HIPCHK(hipMemsetAsync(q, 0, bytes, r.s));
            if (k < n_read) p.in[k] = (const float *)q;
            else p.out[k - n_read] = (float *)q;
        }
        p.n2 = elements / 2;
        const unsigned blocks = (unsigned)((p.n2 + 255) / 256);
        const SF fn = table[n_read - 1][n_write - 1];
        for (int k = 0; k < 2; ++k) hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, r.s, p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(r.a, r.s));
        for (int k = 0; k < iterations; ++k) hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, r.s, p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(r.b, r.s));
        HIPCHK(hipEventSynchronize(r.b));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        *gb_per_s = (double)(n_read + n_write) * (double)bytes * iterations / ((double)ms * 1e-3) / 1e9;
    });
}

int kh_set_timing(kh_ctx *ctx, int32_t enable) {
    return guarded([&] {
        CTX_OK;
        require(enable >= 0, "enable must be >= 0");
        ctx->impl->set_timing(enable);
    });
}

int kh_get_timing(kh_ctx *ctx, double out[2 * KH_TIMING_SLOTS]) {
    return guarded([&] {
        CTX_OK;
        require(out != nullptr, "null argument");
        ctx->impl->get_timing(out);
    });
}

int kh_get_stream(kh_ctx *ctx, void **stream) {
    return guarded([&] {
        CTX_OK;
        require(stream != nullptr, "null argument");
        *stream = ctx->impl->stream();
    });
}

}  // extern "C"
