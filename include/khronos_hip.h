/*
 * khronos_hip.h -- C ABI of libkhronos_hip, the MI355X (gfx950) FDTD time step.
 *
 * The library replaces the reference's per-step device work: the Yee E/H curl
 * updates with the fused PML auxiliary cascade, current-source injection and
 * DFT-monitor accumulation, plus the flux / DFT-norm reductions read at the end
 * of a run.  Host setup (grid volumes, PML profiles, permittivity rasters,
 * source amplitudes, waveforms) stays with the caller, exactly as it does in
 * the reference, and is handed over as plain arrays in the reference's own
 * layouts (column-major, x fastest).
 *
 * Conventions
 *   - every entry point returns 0 on success and a negative KH_E* code on
 *     failure; kh_last_error() then holds a message (thread-local).  Nothing
 *     aborts and no C++ exception crosses the ABI.
 *   - "backend precision" is the simulation's real type (KH_F32 -> float,
 *     KH_F64 -> double).  Complex arrays interleave (re, im) in that type.
 *   - grid indices are the reference's 1-based GridVolume indices
 *     (src/utils.jl:103-214); a volume starting at start_idx with extent dims
 *     covers start_idx .. start_idx + dims - 1 on each axis.
 *   - a context is used by one host thread at a time; distinct contexts are
 *     independent.  The library copies every input array; host buffers are
 *     never retained.
 */
#ifndef KHRONOS_HIP_H
#define KHRONOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KH_ABI_VERSION 8

enum kh_precision { KH_F32 = 0, KH_F64 = 1 };

/* Field component codes (src/DataStructures.jl:24-35).  D and B
 * (KH_DX..KH_BZ) are accepted by kh_add_dft only. */
enum kh_component {
    KH_EX = 0, KH_EY = 1, KH_EZ = 2, KH_HX = 3, KH_HY = 4, KH_HZ = 5,
    KH_DX = 6, KH_DY = 7, KH_DZ = 8, KH_BX = 9, KH_BY = 10, KH_BZ = 11
};

/* KH_MODE_FAST: cells with no PML, source or polarisation term advance E and
 *   H directly (E += eps^-1 * dt * curl H); other cells run the full D/B
 *   cascade.  Results agree with the reference within rounding.
 * KH_MODE_EXACT: every cell keeps D/B and follows the reference CPU path's
 *   operation order (src/Kernels/ReferenceKernels.jl:276-313, :468-512). */
enum kh_mode { KH_MODE_FAST = 0, KH_MODE_EXACT = 1 };

enum kh_error {
    KH_OK = 0,
    KH_E_INVALID = -1,   /* bad argument */
    KH_E_NOMEM = -2,     /* device allocation failed */
    KH_E_DEVICE = -3,    /* HIP runtime error */
    KH_E_UNSUPPORTED = -4,
    KH_E_STATE = -5      /* call out of order (e.g. source added after stepping) */
};

typedef struct kh_ctx kh_ctx;

/* Everything kh_create needs.  Pointers may be NULL where noted. */
typedef struct kh_sim_desc {
    int32_t nx, ny, nz;        /* cells per axis (the reference's Nx, Ny, Nz) */
    int32_t precision;         /* enum kh_precision */
    int32_t mode;              /* enum kh_mode */
    int32_t dft_accum_f64;     /* 1: DFT accumulators in float64 (default form);
                                  0: in the backend precision, as the reference */
    double dx, dy, dz, dt;     /* backend-precision values (exact in double) */
    /* PML profiles, length 2*n+1 each, backend precision
     * (src/Boundaries.jl:64-72, :112-163).  NULL = no PML arrays at all.
     * 2D simulations (the reference's Nz = 0) pass nz = 1, dz = +inf and NULL
     * for the two z profiles, which the reference does not create (:130-162). */
    const void *sigma_b[3];
    const void *sigma_d[3];
    /* Per-voxel inverse permittivity / permeability of Ex,Ey,Ez / Hx,Hy,Hz in
     * the reference geometry layout: component voxel count, no ghosts
     * (src/Geometry.jl:459-473).  NULL = use the scalar below. */
    const void *eps_inv[3];
    const void *mu_inv[3];
    double eps_inv_scalar;     /* src/Geometry.jl:662 */
    double mu_inv_scalar;
    double source_cutoff_time; /* last_source_time (src/Sources/TimeSources.jl:46);
                                  +inf for sources without cutoff */
    /* Periodic / Bloch boundaries (src/Chunking.jl:1725-1768).  complex_fields
     * = 1 when any side is Bloch (src/Fields.jl:145-159): every field is then
     * complex; kh_read_field / kh_write_field / kh_read_pole / kh_write_pole
     * exchange interleaved (re, im) pairs and current sources still add only
     * Re(f * amplitude) (src/Sources/Sources.jl:346-357).  periodic[a] = 1 when
     * axis a wraps: after each half step the updated group gets
     * ghost 0 <- plane N * bloch_rev[a] and ghost N+1 <- plane 1 * bloch_fwd[a]
     * (complex (re, im) factors exp(-i k L) and exp(+i k L); 1 for Periodic). */
    int32_t complex_fields;
    int32_t periodic[3];
    double bloch_fwd[3][2];
    double bloch_rev[3][2];
    /* Material conductivity sigma_D of Ex,Ey,Ez (0..2) and sigma_B of
     * Hx,Hy,Hz (3..5) per voxel, reference geometry layout, absorber profiles
     * already added (src/Geometry.jl:467-476, :117-118, :708-789).  A component
     * with one runs the C stage of the cascade in every cell, sigma~ =
     * 0.5 * (dt * sigma) (src/Kernels/Helpers.jl:39-70, :278-279).  Needs the
     * six PML profiles above (zeros where there is no PML).  NULL = none. */
    const void *sigma_mat[6];
} kh_sim_desc;

/* Message describing the last failure on this thread ("" if none). */
const char *kh_last_error(void);

/* KH_ABI_VERSION. */
int kh_version(void);

/* Device bytes kh_create would allocate for `desc` (fields, auxiliaries,
 * material arrays); sources and monitors add their own arrays.
 * Replaces: estimate_memory (src/Memory.jl:120-251). */
int kh_mem_footprint(const kh_sim_desc *desc, size_t *bytes);

/* Allocate a simulation on the current device and upload its materials.
 * Replaces: init_fields / create_all_chunks / init_boundaries uploads
 * (src/Simulation.jl:92-283, src/Fields.jl:131-142). */
int kh_create(const kh_sim_desc *desc, kh_ctx **out);

/* Free every device array of the context. */
int kh_destroy(kh_ctx *ctx);

/* Register a current source: S[start + i] += Re(f(t) * amplitude[i]) on the
 * component's half step while sources are active.  `amplitude` is complex,
 * backend precision, dims[0]*dims[1]*dims[2] values, column-major.  Any
 * number of sources; those of a half step are added in registration order.
 * Replaces: assemble_sources upload + update_source! (src/Sources/Sources.jl:44-131,
 * :346-357). */
int kh_add_source(kh_ctx *ctx, int32_t component, const int32_t start_idx[3],
                  const int32_t dims[3], const void *amplitude, int32_t *source_id);

/* Waveform of a source: values[2*n], values[2*n+1] = Re, Im of f(t_n) for
 * timestep n = 0..n_steps-1, where t_n is the time the reference evaluates
 * the source at on that step (t for H sources, t + dt/2 for E sources;
 * src/Kernels/Kernels.jl:64-74).  Steps past the table inject zero.
 * Replaces: eval_time_source per step (src/Sources/TimeSources.jl:123-132). */
int kh_set_waveform(kh_ctx *ctx, int32_t source_id, int64_t n_steps, const double *values);

/* Register a DFT monitor on one component: every `decimation`-th step,
 * F[i,k] += dt * exp(i*2*pi*freqs[k]*t) * field[start + i]
 * (H and B at t, E and D at t + dt/2).  D and B monitors record the D/B
 * arrays in KH_MODE_EXACT; in KH_MODE_FAST, which keeps no D/B, the value is
 * rebuilt from the carried stage, (W where the own-axis PML sigma is
 * non-zero, else E or H) / eps^-1 (mu^-1) -- exact outside dispersive
 * materials and source cells.  In KH_MODE_FAST the E update applies the
 * update of up to four small E monitors itself (real fields, float64
 * accumulators, no dispersive/chi3 media or periodic axis, at most 8192
 * cell-frequencies inside cells 1..N, no E source on a monitor cell of its
 * component; KH_DFT_LAUNCH=1: never); the result is the same bit for bit.
 * Replaces: init_monitors(DFTMonitor) and
 * update_dft_monitor! (src/Monitors/Monitors.jl:202-290, :361-379); the
 * design-region D monitors of install_design_region_monitors!
 * (src/Adjoint/DesignRegion.jl:290-312). */
int kh_add_dft(kh_ctx *ctx, int32_t component, const int32_t start_idx[3],
               const int32_t dims[3], int32_t n_freq, const double *freqs,
               int32_t decimation, int32_t *dft_id);

/* Replace the per-voxel inverse permittivity of component 0..2 (reference
 * geometry layout, as in kh_sim_desc.eps_inv); the context must have been
 * created with per-voxel permittivity.  Takes effect from the next step;
 * fields are untouched.  Replaces: update_design! / _write_design_to_sim!
 * (src/Adjoint/DesignRegion.jl:211-288). */
int kh_set_eps_inv(kh_ctx *ctx, int32_t component, const void *eps_inv);

/* Kerr coefficient per cell, (nx, ny, nz) column-major, backend precision,
 * painted on the cell centres (src/Geometry.jl:610-636).  After every E update
 * each cell with chi3 != 0 gets E <- E / (1 + chi3 (Ex^2 + Ey^2 + Ez^2)) at the
 * same grid index, before the polarization update and the E monitors.
 * Before the first kh_step.  Replaces: _chi3_correction_kernel! /
 * step_chi3_correction! (src/Kernels/Dispersive.jl:127-173). */
int kh_set_chi3(kh_ctx *ctx, const void *chi3);

/* One Lorentz or Drude pole.  coeffs = {gamma1_inv, gamma1, omega0_dt_sq,
 * sigma_omega0_dt_sq, drude_coeff} as compute_ade_coefficients returns them
 * (src/Susceptibility.jl:67-85), is_drude = (omega_0 == 0); sigma is the
 * oscillator strength per cell, (nx, ny, nz) column-major, backend precision,
 * already zeroed inside the PML (src/Geometry.jl:1180-1222).  The library
 * keeps P and P_prev over the bounding box of sigma != 0; after every E update
 * (and chi3) P <- gamma1_inv (c P - gamma1 P_prev + k sigma E), and the E
 * update subtracts the sum of all poles' P.  The chi1 correction of the
 * inverse permittivity is the caller's (src/Geometry.jl:1224-1350), as in the
 * reference.  Before the first kh_step.  Replaces: init_polarization!
 * uploads and step_polarization! (src/Geometry.jl:1136-1355,
 * src/Kernels/Dispersive.jl:25-117, :186-228). */
int kh_add_pole(kh_ctx *ctx, const double coeffs[5], int32_t is_drude, const void *sigma,
                int32_t *pole_id);

/* P (which = 0) or P_prev (which = 1) of component 0..2 of a pole, in the
 * reference's polarization layout (nx+2, ny+2, nz+2), column-major, backend
 * precision; zero outside the pole's box.  Synchronous.  Replaces:
 * Array(pole.Px) etc. (src/DataStructures.jl:288-304). */
int kh_read_pole(kh_ctx *ctx, int32_t pole_id, int32_t which, int32_t component, void *host,
                 size_t bytes);
int kh_write_pole(kh_ctx *ctx, int32_t pole_id, int32_t which, int32_t component,
                  const void *host, size_t bytes);

/* Advance n_steps time steps in the reference's sub-step order
 * (H sources, H update, H monitors, E sources, E update, E monitors).
 * Asynchronous with respect to the host.  Replaces: step! (src/Kernels/Kernels.jl:20-88). */
int kh_step(kh_ctx *ctx, int64_t n_steps);

/* Steps per captured graph (0 = launch every kernel of every step).  kh_step
 * captures blocks of steps as hipGraphs and replays them; the launches of a
 * step carry no per-step arguments (the timestep, waveform values and DFT
 * gates are read on the device), so sources and monitors stay inside the
 * graph.  The main block is steps_per_graph rounded up to the DFT period
 * (the lcm of the monitors' decimations) and starts on a multiple of it;
 * shorter or unaligned stretches replay blocks of steps_per_graph steps
 * captured for their start phase within the period (the last few phases are
 * kept); the rest runs eagerly.  Graphs are not used while kernel timing is
 * on.  Default 16, or the environment's KH_GRAPH_STEPS.
 * Replaces: _try_capture_graphs! and the graph replay branch of step!
 * (src/Kernels/Kernels.jl:40-61, :99-145; KHRONOS_CUDA_GRAPHS=0 there). */
int kh_set_graph(kh_ctx *ctx, int32_t steps_per_graph);

/* Capture now the graphs that kh_step(ctx, n_steps) would replay from the
 * current state, without running any step, so that a timed kh_step replays
 * graphs captured beforehand.  No counterpart in the reference (its graphs
 * are captured on first use, src/Kernels/Kernels.jl:40-61). */
int kh_prepare_graphs(kh_ctx *ctx, int64_t n_steps);

/* What kh_step ran since the last call (then cleared): counts[0] graph
 * replays, [1] steps inside those replays, [2] steps launched eagerly,
 * [3] DFT kernel launches (inside replays and eager; monitors the E
 * update applies itself are not counted).  For measurement;
 * no counterpart in the reference. */
int kh_get_launch_counts(kh_ctx *ctx, int64_t counts[4]);

/* Whether all queued work of the context has finished (*idle = 1), without
 * waiting; a scheduler serving many contexts polls this to keep each one's
 * stream fed.  Replaces: the task fetch of run_batch_concurrent
 * (src/Batch.jl:183-263). */
int kh_query(kh_ctx *ctx, int32_t *idle);

/* Wait for all queued work of the context. */
int kh_sync(kh_ctx *ctx);

/* Current timestep counter (src/DataStructures.jl:750). */
int kh_get_timestep(kh_ctx *ctx, int64_t *timestep);

/* Whether sources are still injected (src/DataStructures.jl:751). */
int kh_sources_active(kh_ctx *ctx, int32_t *active);

/* Copy one field component to/from `host`, in the reference's array layout:
 * (component voxel count + 2) per axis, ghosts included, column-major,
 * backend precision.  `bytes` must equal that size.  Synchronous. */
int kh_read_field(kh_ctx *ctx, int32_t component, void *host, size_t bytes);
int kh_write_field(kh_ctx *ctx, int32_t component, const void *host, size_t bytes);

/* Copy a DFT accumulator to `host` as complex128, shape (dims..., n_freq),
 * column-major.  Synchronous.  Replaces: Array(monitor.fields). */
int kh_read_dft(kh_ctx *ctx, int32_t dft_id, double *host, size_t bytes);

/* Poynting flux of a planar monitor from its four tangential DFTs
 * (e1, e2, h1, h2), normal_axis 1..3, area element dA; out[n_freq] in
 * float64.  Reduction on the device.  Replaces: get_flux
 * (src/Monitors/FluxMonitor.jl:92-156). */
int kh_flux(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2,
            int32_t normal_axis, double dA, double *out);

/* Mode overlap integrals of a planar mode monitor from its four tangential
 * DFTs (e1, e2, h1, h2 as for kh_flux): `mode` holds the mode's tangential
 * E (m1, m2) and H (n1, n2) sampled where the normal-averaged DFT samples
 * lie, complex (re, im) float64, index ((q * n_freq + k) * n2 + i2) * n1 + i1
 * for q = m1, m2, n1, n2 and i1 < n1, i2 < n2 along the two tangential axes
 * in increasing axis order.  out[4 * k ..] = sum (s + t) dA and sum (s - t) dA
 * as (re, im), with s = (E_sim x H_mode*) . n and t = (E_mode* x H_sim) . n
 * in kh_flux's component order; a+- = these / (4 P_mode).  Reduction on the
 * device.  Replaces: compute_mode_amplitudes
 * (src/Monitors/ModeMonitor.jl:345-507). */
int kh_mode_overlap(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis,
                    int32_t n1, int32_t n2, const double *mode, double dA, double *out);

/* Near-to-far projection of a planar surface from its four tangential DFTs
 * (e1, e2, h1, h2 as for kh_flux): the full dyadic Green's function (1/r,
 * 1/r^2, 1/r^3 terms) of the equivalent currents J, M of every sample of the
 * n1 x n2 plane (i1 along the lower tangential axis), summed at each
 * observation point, in float64.  normal_sign is +1 or -1 (outward normal
 * direction); base[4][3] is the position of sample (0, 0) of e1, e2, h1, h2;
 * d1, d2 the tangential spacings (lower axis first) and dA = d1 * d2;
 * freqs[n_freq] must be the DFTs' frequencies; obs[n_obs][3] the
 * observation points.  out receives complex (re, im) (Ex, Ey, Ez, Hx, Hy, Hz)
 * at index (k * 6 + j) * n_obs + o: the column-major (n_obs, 6, n_freq)
 * array of the reference.  Synchronous.  Replaces: _compute_far_field_gpu
 * and near2far_kernel_{x,y,z}! (src/Monitors/Near2Far.jl:103-246, :380-550),
 * green3d! (:40-94). */
int kh_near2far(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis,
                double normal_sign, const double *base, double d1, double d2, double dA, int32_t n1,
                int32_t n2, const double *freqs, int32_t n_freq, double medium_eps, double medium_mu,
                const double *obs, int32_t n_obs, double *out);

/* Spectral samples of the four tangential DFT planes of a surface (as for
 * kh_near2far) at frequency index freq_index, for the layered projection:
 * out[8 * b + 2 * q ..] = sum_{i1 < n1, i2 < n2} F_q(i1, i2)
 * exp(-2 pi i (m1 i1 / pad1 + m2 i2 / pad2)) as (re, im), q = e1, e2, h1, h2,
 * for bin b = (m1, m2) = bins[2 * b], bins[2 * b + 1] (integer frequencies,
 * any sign).  These are the bins of the zero-padded pad1 x pad2 FFT the
 * reference interpolates.  Synchronous.  Replaces: the padded FFTs of
 * compute_far_field_layered (src/Monitors/Near2Far.jl:805-840). */
int kh_near2far_spectrum(kh_ctx *ctx, int32_t e1, int32_t e2, int32_t h1, int32_t h2, int32_t normal_axis,
                         int32_t n1, int32_t n2, int32_t freq_index, int32_t pad1, int32_t pad2,
                         const int32_t *bins, int32_t n_bins, double *out);

/* Continuous-batching scheduler over contexts (one GPU).  Replaces:
 * run_batch_concurrent and plan_batch (src/Batch.jl:142-263, :22-69), whose
 * waves of tasks on separate streams become a rolling admission.
 *   kh_server_admit: reserve a slot and `footprint` bytes of the budget
 *     (*admitted = 1) if fewer than max_active requests are in flight and the
 *     budget holds it, or nothing else is in flight; then build the context.
 *   kh_server_submit: hand over the admitted request's context for n_steps
 *     steps; the caller keeps ownership and must not touch the context until
 *     the request is reported done.  kh_server_cancel returns an admission
 *     that will not be submitted.
 *   kh_server_poll: without blocking, keeps two chunks of chunk_steps steps
 *     queued on every submitted context (kh_step on its own stream; the next
 *     chunk waits while the current one runs, so a context does not drain
 *     between polls) and retires those with no steps left and an idle
 *     stream; done_ids receives up to cap finished request ids in
 *     completion order, *n_done their count, *n_inflight (may be NULL) the
 *     requests admitted and not yet reported.
 * admit/submit/cancel may be called from other threads than poll; each
 * context is stepped from the polling thread only. */
typedef struct kh_server kh_server;
int kh_server_create(double budget_bytes, int32_t max_active, int32_t chunk_steps, kh_server **out);
int kh_server_destroy(kh_server *server);
int kh_server_admit(kh_server *server, double footprint, int32_t *admitted);
int kh_server_submit(kh_server *server, kh_ctx *ctx, int64_t n_steps, double footprint, int64_t request_id);
int kh_server_cancel(kh_server *server, double footprint);
int kh_server_poll(kh_server *server, int64_t *done_ids, int32_t cap, int32_t *n_done, int32_t *n_inflight);

/* Geometry objects for kh_raster_component.  shape: KH_SHAPE_BOX (size =
 * full extents), KH_SHAPE_SPHERE (size[0] = radius), KH_SHAPE_CYLINDER
 * (size[0] = radius, size[1] = height, axis 0..2), KH_SHAPE_PRISM (a convex
 * polygon of n_vertices (x, y) pairs about center, counter-clockwise, in
 * the x-y plane, extruded over size[0] along z: the prisms of the GDS
 * import, src/GDS.jl:954-1002).  value[c]: relative permittivity of Ex, Ey,
 * Ez (c = 0..2) or permeability of Hx, Hy, Hz (3..5) inside the object; NaN
 * where the object does not set it.  vertices is read only for prisms. */
enum kh_shape { KH_SHAPE_BOX = 0, KH_SHAPE_SPHERE = 1, KH_SHAPE_CYLINDER = 2, KH_SHAPE_PRISM = 3 };
typedef struct kh_raster_object {
    int32_t shape;
    int32_t axis;
    double center[3];
    double size[3];
    double value[6];
    int32_t n_vertices;
    int32_t reserved;
    const double *vertices;
} kh_raster_object;

/* Rasterise one material component on the current device: out (precision
 * type, cnt[0] * cnt[1] * cnt[2] values, column-major, the reference geometry
 * layout) receives 1 / value of the first object in list order whose
 * bounding box and shape hold the component's grid point (xs[i], ys[j],
 * zs[k]) and that sets the component, else 1 (the reference paints objects
 * last to first).  smoothing (electric components only): 0 none,
 * 1 volume averaging, 2 anisotropic (Farjadpour); half[3] = half the cell
 * size.  A voxel differing (relative 1e-6) from a face neighbour takes the
 * effective value of the plane through the nearest shape surface.
 * Synchronous; allocates and frees its own device memory.  Replaces: the
 * painting of init_geometry and smooth_eps!
 * (src/Geometry.jl:450-663, :792-955). */
int kh_raster_component(int32_t precision, int32_t component, const int32_t cnt[3], const double *xs,
                        const double *ys, const double *zs, const kh_raster_object *objs, int32_t n_objs,
                        int32_t smoothing, const double half[3], void *out);

/* Bytes of device staging available to kh_raster_component_noalloc. */
#define KH_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* Allocation-free variant of kh_raster_component.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (a
 * serving loop preparing requests beside running streams, where hipMalloc
 * and hipFree would synchronise the device; code under a harness that forbids
 * allocation).  Same arguments and values as kh_raster_component, except:
 * out_device is device memory of cnt[0] * cnt[1] * cnt[2] values of the
 * precision type, and the work runs on `stream` (a hipStream_t; NULL for the
 * null stream); the call returns once that stream has finished it.  Staging
 * (objects, coordinates, prism vertices and, with smoothing, the unsmoothed
 * raster) lives in a device array of KH_NOALLOC_SCRATCH_BYTES (64 MiB) that
 * each GPU holds until the process ends; a call that needs more fails with
 * KH_E_INVALID.  Calls that overlap on one GPU share that array and race. */
int kh_raster_component_noalloc(int32_t precision, int32_t component, const int32_t cnt[3], const double *xs,
                                const double *ys, const double *zs, const kh_raster_object *objs, int32_t n_objs,
                                int32_t smoothing, const double half[3], void *out_device, void *stream);

/* sqrt(sum |F|^2) of DFT monitor `dft_id` in float64, reduced on the device;
 * dft_id = -1 gives the sum of that over all monitors.
 * Replaces: dft_fields_norm (src/Simulation.jl:534-539). */
int kh_dft_norm(kh_ctx *ctx, int32_t dft_id, double *out);

/* out[i] = sqrt(sum |F|^2) of DFT monitor i, for all n_out monitors of the
 * context (n_out must equal their number), with one synchronisation per 64
 * monitors; the same values as kh_dft_norm monitor by monitor.
 * Replaces: the per-monitor dft_fields_norm calls of the
 * stop_when_dft_decayed callback (src/Simulation.jl:437-481). */
int kh_dft_norms(kh_ctx *ctx, int32_t n_out, double *out);

/* Zero fields, auxiliaries, DFT accumulators and the timestep; re-arm sources.
 * Replaces: reset_fields! (src/Simulation.jl:571-637). */
int kh_reset(kh_ctx *ctx);

/* Forward-field checkpoints for the time-reversed adjoint: every `every`
 * steps of the run (from timestep 0) the whole field state is copied in
 * device memory; 0 turns them off and frees them.  Only before the first
 * step or right after kh_reset (KH_E_STATE otherwise).  Writing fields,
 * permittivity or poles afterwards invalidates the checkpoints taken so far.
 * Serves the reverse rule of simulate (src/Adjoint/ChainRulesIntegration.jl:
 * 38-60), whose gradient the reference gets from a second, frequency-domain
 * run instead (src/Adjoint/OptimizationProblem.jl:169-330). */
int kh_set_checkpoints(kh_ctx *ctx, int32_t every);

/* Gradient dJ/d(eps^-1) of a real objective J(F) of the listed DFT monitors
 * (E or H components) after the steps run since timestep 0, by the discrete
 * transpose of those steps: each checkpoint interval is re-run from its
 * checkpoint (monitors off) and walked backwards through the transposed H/E
 * half steps, CPML stages included.
 * `cotangent`: per listed monitor in order, dJ/dRe(F) + i dJ/dIm(F) as
 * (re, im) pairs in kh_read_dft's layout.  Design boxes, one per E component
 * (a = 0..2): first grid index box_lo[3a..3a+2] (1-based, cells 1..N) and
 * extent box_n[3a..3a+2] (any 0: component skipped).  `grad`: float64, the
 * boxes in order, each
 * dense with x fastest.  Real fields without dispersive, chi3 or conductive
 * media and without periodic axes.  The fields, timestep and monitors are
 * those of the forward run again on return (the fields restored in the
 * order of the context's stream).  Unless graphs are off (kh_set_graph 0)
 * or kernel timing is on, each full interval of up to 1024 steps (re-run,
 * snapshots, walk back) replays graphs captured for the call; the result is
 * the same bit for bit.  In FAST mode the re-run's E
 * update writes the design-box snapshots itself unless an E source writes a
 * box cell of its component (KH_SNAP_LAUNCH=1: always a separate copy).
 * The walk back of each interval runs on a second stream of the context
 * while the next earlier interval is re-run (KH_SWEEP_STREAMS=0: one
 * stream); the call returns after both have finished.
 * Replaces: the pullback of simulate (src/Adjoint/ChainRulesIntegration.jl:
 * 38-60) with calculate_gradient! (src/Adjoint/GradientKernel.jl:35-109),
 * as dJ/d(eps^-1) per voxel before the design-grid contraction. */
int kh_adjoint_gradient(kh_ctx *ctx, int32_t n_monitors, const int32_t *monitor_ids, const double *cotangent,
                        const int32_t box_lo[9], const int32_t box_n[9], double *grad);

/* Kernel timing for measurement: while enabled (enable > 0), every update
 * and DFT launch stamps a HIP event pair at its kernel's start and end on the
 * context's stream (hipExtLaunchKernel: the dispatch gap before the kernel is
 * not counted); `enable`
 * event pairs are created up front so that timing allocates nothing while it
 * runs.  kh_get_timing writes {total ms, launch count} for each of the
 * KH_TIMING_SLOTS kernel classes -- H inner box, H PML/full cascade, E inner
 * box, E PML/full cascade, DFT, material (chi3 + poles) -- and clears the
 * counters. */
#define KH_TIMING_SLOTS 6
int kh_set_timing(kh_ctx *ctx, int32_t enable);
int kh_get_timing(kh_ctx *ctx, double out[2 * KH_TIMING_SLOTS]);

/* The hipStream_t the context launches on (for callers that order their own
 * work after it). */
int kh_get_stream(kh_ctx *ctx, void **stream);

/* Streaming rate of a plain one-pass kernel on the current device, the
 * practical HBM ceiling for a read/write mix: n_read float32 arrays of
 * `elements` values read and n_write written (two values per lane,
 * nontemporal, 256-lane blocks in linear order), `iterations` launches timed
 * with HIP events after two untimed ones; *gb_per_s = bytes moved / time.
 * n_read 1..6, n_write 1..3, elements even and >= 512; the arrays are
 * allocated and freed by the call.  No counterpart in the reference
 * (measurement only). */
int kh_stream_rate(int32_t n_read, int32_t n_write, int64_t elements, int32_t iterations, double *gb_per_s);

#ifdef __cplusplus
}
#endif

#endif /* KHRONOS_HIP_H */
