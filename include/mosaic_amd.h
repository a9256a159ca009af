/*
 * mosaic_amd -- C ABI of the MI355X (gfx950) kinematic-scattering and
 * propagation path.
 *
 * Conventions shared by every entry point:
 *   - plain pointers and sizes only; "device" pointers are HIP device
 *     memory, "host" pointers are ordinary host memory;
 *   - every call is ordered on the HIP stream it is given (0 = the null
 *     stream) and returns an int status (MOSAIC_OK or an error below);
 *   - scratch memory comes from the caller as a workspace whose size a
 *     *_workspace query returns; the library keeps only hipFFT plans (per
 *     device and size) and the allocation-free entry point's fixed scratch;
 *   - complex arrays are interleaved (re, im) pairs.
 *
 * Pixel layout follows the reference: p = iz * Ny + iy, fields (Nz, Ny)
 * row-major (Beam.py:704, 719-720).
 */
#ifndef MOSAIC_AMD_H
#define MOSAIC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOSAIC_OK 0
#define MOSAIC_ERR_ARGS 1          /* a size or pointer is invalid */
#define MOSAIC_ERR_HIP 2           /* a HIP runtime call failed */
#define MOSAIC_ERR_UNSUPPORTED 3   /* declined: outside the fast path's error bound */
#define MOSAIC_ERR_WORKSPACE 4     /* workspace smaller than the query returned */
#define MOSAIC_ERR_TOO_LARGE 5     /* beyond the fixed scratch of an allocation-free call */

/* Opaque stream handle (a hipStream_t). */
typedef void* mosaic_stream_t;

/* Per-species constants of the kinematic sum. */
typedef struct {
    float wk[11];     /* Waasmaier-Kirfel [a1..a5, c, b1..b5] (Beam.py:503-527) */
    float f0_zero;    /* f0(0) = c + sum a_i (Beam.py:600-615) */
    float fp;         /* f'  at the beam energy (Beam.py:570-598) */
    float fpp;        /* f'' at the beam energy */
} mosaic_species_t;

/* Flags and beam constants of the kinematic sum. */
typedef struct {
    double k;               /* wavenumber 2 pi / lambda [rad/m] */
    int remove_forward;     /* subtract f0(0)+f'+if'' below the pixel's Q cut */
    int apply_pol;          /* scale by sqrt(P) */
    int apply_decay;        /* scale by R0 / r */
    int path;               /* 0 automatic, 1 per-tile factors, 2 row recurrence,
                               3 matrix-core patches, 4 moment expansion,
                               5 exact per pair, 6 probe (few pixels),
                               7 moment expansion on the matrix cores */
    double pol_deg;         /* degree of linear polarization p */
    double pol_e[3];        /* its unit vector (lab) */
    int analyser_kind;      /* 0 off, 1 top-hat, 2 Darwin roll-off (Beam.py:2350-2378);
                               any other than 0 takes the exact per-pair path */
    int gm_pixel_eval;      /* path 7: evaluate every (unit, patch) pixel by pixel
                               instead of by the separable phase model (the
                               model's own fallback; slower, same bar) */
    double analyser_axis[3];       /* analyser axis (lab); the reference uses the
                                      detector's centre pixel (Beam.py:6128-6132) */
    double analyser_accept_rad;    /* top-hat half-angle */
    double analyser_darwin_hw_rad; /* Darwin half-width; <= 0: weight 1 */
} mosaic_kin_opts_t;

/*
 * Workspace bytes for mosaic_kin_stage + mosaic_kin_sum on a sample of
 * `atom_count` atoms split by species as `species_counts[n_species]`, for
 * `n_frames` poses of an Ny x Nz detector.  Limits (MOSAIC_ERR_ARGS beyond
 * them, here and in every staged-sum call): fewer than 2^32 atoms, at most
 * 16 species, 65535 poses per call, 2^30 pixels and 2^31 - 1 pixel-poses.
 */
int mosaic_kin_workspace(int64_t atom_count, int n_species,
                         const int64_t* species_counts /* host */,
                         int n_frames, int Ny, int Nz, size_t* bytes);

/*
 * Stage a sample into the workspace: group the atoms by species, cut each
 * species into 256-atom tiles by a k-d bisection (compact boxes of equal
 * count), and store each tile as an FP64 origin plus FP32 offsets in
 * wavelengths with the incident amplitude.  Replaces the reference's
 * per-species partition, Morton order and per-chunk origins
 * (_partition/_stage, Beam.py:9265-9319; _morton_order_gpu/_chunk_origins_gpu,
 * Beam.py:1205-1300).
 *
 *   pos      device (N, 3) float32, sample frame, Angstrom
 *   species  device (N,) uint8 species codes (< n_species)
 *   amp      device (N,) complex64 incident amplitude, or NULL for 1
 *   stage_stats  host out [2]: largest tile radius in wavelengths, largest
 *                tile-origin distance from the sample origin in metres, or
 *                NULL to leave them in the workspace (mosaic_kin_stats,
 *                mosaic_kin_sum read them there)
 * The call synchronizes the stream to return stage_stats; with NULL it
 * returns without waiting.
 */
int mosaic_kin_stage(int64_t atom_count, const float* pos, const uint8_t* species,
                     const float* amp, int n_species,
                     const int64_t* species_counts /* host */, double k,
                     int n_frames, int Ny, int Nz, void* workspace,
                     size_t workspace_bytes, double* stage_stats,
                     mosaic_stream_t stream);

/*
 * Re-stage moved atoms without sorting: rebuild the tiles of a workspace that
 * mosaic_kin_stage filled for the same atoms (same count, species codes and
 * order), from new positions and amplitudes, keeping the tile assignment.
 * For displacements small against a tile (strain fields, defects, thermal
 * draws, fitting steps) the tiles stay compact; stage_stats reports their
 * radius, which the sum's bounds check as usual.  The radii of the pairs and
 * quads of tiles that the moment sum groups into units are measured again
 * from the new positions.  The counterpart of the
 * reference's per-step re-staging (reload, Morton order and origins per scan
 * step, Experiment.py:305-310; Beam.py:1205-1300), minus the sort.
 * Synchronizes the stream to return stage_stats.
 */
int mosaic_kin_restage(int64_t atom_count, const float* pos, const float* amp,
                       int n_species, const int64_t* species_counts /* host */, double k,
                       int n_frames, int Ny, int Nz, void* workspace,
                       size_t workspace_bytes, double* stage_stats,
                       mosaic_stream_t stream);

/*
 * Pixel prologue of the kinematic sum, once per detector geometry, species
 * table and flags: writes into the head of the workspace, per pixel,
 * rho0 = |T_p|, frac(rho0 / lambda) in double-double, the forward-removal
 * Q cut (Beam.py:718-746), the unit vector and 1/(2 rho0), and per species
 * the factor (f0 + f' + i f'') r_e sqrt(P) in the pixel's direction.
 * Returns pixel_geom (host out [8]): the smallest pixel distance [m], the
 * largest first and second differences of the pixel direction along a row,
 * the largest deviation of the pixel direction from the separable per-patch
 * quadratic model of the far-field sums, the largest second and first
 * differences along a column, the largest deviation from the model with the
 * cross term a b g_yz (the moment expansion's), and the largest |g_yz|; the
 * call synchronizes the stream for them.  With pixel_geom NULL they stay in
 * the workspace and the call returns without waiting.  Must precede
 * mosaic_kin_sum on the same workspace.  `coords` is device (3, Ny*Nz)
 * float64 metres.
 */
int mosaic_kin_pixels(int n_species, const mosaic_species_t* species_tab /* host */,
                      const mosaic_kin_opts_t* opts /* host */, int Ny, int Nz,
                      const double* coords, void* workspace, size_t workspace_bytes,
                      double* pixel_geom, mosaic_stream_t stream);

/*
 * Kinematic sum of a staged sample for `n_frames` sample poses
 *   lab = rot[f] @ sample + trans[f]   (rot row-major float32 3x3, trans Angstrom)
 * onto pixels `coords` (device (3, Ny*Nz) float64 metres), accumulated into
 * `out` (device (n_frames, Nz, Ny) complex64, or complex128 when
 * out_is_c128).  Replaces the per-chunk kernel launches of interact_beam_gpu
 * (Beam.py:6279-6377): the fast path behind dispatch (mosaic_fast,
 * Beam.py:8246-8632, 9336-9496) and its fallback, the general kernel
 * (interaction_kernal, Beam.py:1912-2412), with compute_scattering_cffi's
 * per-pair semantics (Beam.py:663-816).  It works on a staged sample (sample
 * frame, species codes, incident amplitudes), not on the reference's lab
 * positions in metres and per-atom table rows: mosaic_scatter below takes
 * those.  Five kernels:
 *   - exact per pair (path 5): no expansion; the distance, form factor,
 *     polarization, decay and analyser weight of every atom-pixel pair, with
 *     an FP64 (tile, pixel) origin phase and FP32 offsets;
 *   - per-tile factors (1): exact distance per tile and pixel, form factor,
 *     polarization and decay once per (tile, pixel), optional second-order
 *     offset term; it needs the tile radius below 1e-6 of the pixel distance;
 *   - and, in the far field, a row recurrence (2: per-pixel factors, far-field
 *     origin phase, one sincos per atom per KP pixels), a matrix-core form over
 *     32 x 32 pixel patches (3: the tile sum as a complex product of per-row and
 *     per-column phasors, FP16-split MFMA with FP32 accumulation) and a moment
 *     expansion over the same patches (4: the sum over a unit of 1, 2 or 4
 *     consecutive tiles as a polynomial of degree D = 5 or 6 in the pixel
 *     offsets, from (D+1)(D+2)/2 moments of the atoms' small-angle phasors;
 *     units of more than one tile are used where their radius, measured by
 *     mosaic_kin_stage, keeps every error bound); the same expansion with its
 *     per-atom work on the matrix cores (7: the degree-6 moments of 64
 *     patches as one FP16-split matrix product over each unit's atoms, with
 *     each patch's gradients expanded to second order about its region's
 *     mean, then evaluated at the pixels from a moment buffer in the
 *     workspace); and a probe form for a few scattered pixels (6: one lane
 *     per (frame, pixel), the recurrence's far-field model with the exact
 *     direction at each pixel, one sincos per pair).
 * opts->path = 0 picks the matrix-core moment sum (from 64 patches up),
 * else the moment expansion, else the matrix-core form, else the
 * recurrence, else (at most 16384 pixel-poses) the probe form, else the
 * per-tile form, whenever their error bounds hold, and the exact kernel
 * otherwise; an analyser always takes the exact kernel.
 * A forced path (1-4, 6, 7) returns MOSAIC_ERR_UNSUPPORTED (nothing
 * written) when its bounds fail.
 * `stage_stats` and `pixel_geom` are the outputs of mosaic_kin_stage and
 * mosaic_kin_pixels, or NULL where those calls left them in the workspace:
 * the call then reads them there and synchronizes the stream once (one wait
 * for both, in place of one in each of those calls); `path_used` (host out, may be NULL) reports 11/12
 * (per-tile, first/second order), 24/28 (row recurrence over 4/8 pixels),
 * 32 (matrix-core), 45/46 (moment expansion of degree 5/6), 76 (the same on
 * the matrix cores), 60 (probe), plus 100 for
 * the far-field forms with the atoms' second-order phase, or 50 + the
 * analyser kind (exact).  Stream-ordered, no host synchronization when both
 * statistics are given.
 */
int mosaic_kin_sum(int64_t atom_count, int n_species,
                   const int64_t* species_counts /* host */,
                   const mosaic_species_t* species_tab /* host [n_species] */,
                   const mosaic_kin_opts_t* opts /* host */,
                   int n_frames, const float* rot /* host */,
                   const float* trans /* host */, int Ny, int Nz,
                   const double* coords, void* out, int out_is_c128,
                   const double* stage_stats /* host [2] */,
                   const double* pixel_geom /* host [8] */, void* workspace,
                   size_t workspace_bytes, int* path_used, mosaic_stream_t stream);

/*
 * The statistics that mosaic_kin_stage and mosaic_kin_pixels left in a
 * workspace when called with NULL outputs: stage_stats (host out [2], or
 * NULL) as mosaic_kin_stage returns them (zeros for an empty sample), and
 * pixel_geom (host out [8], or NULL) as mosaic_kin_pixels returns them.
 * Synchronizes the stream.  No counterpart in the reference: its staging
 * and pixel tables return to Python at once (Beam.py:9265-9319, 9241-9262);
 * here they may stay on the device until the sum needs them.
 */
int mosaic_kin_stats(int64_t atom_count, int n_species,
                     const int64_t* species_counts /* host */, int n_frames, int Ny, int Nz,
                     void* workspace, size_t workspace_bytes, double* stage_stats,
                     double* pixel_geom, mosaic_stream_t stream);

/*
 * Launch times of the kernels of mosaic_kin_sum calls on the calling
 * thread's current device, for measurement.  Switching timing on (enable = 1
 * after 0) starts a fresh record: the matrix-core moment sum (path 7) then
 * records HIP events on its stream around its moments and its evaluation
 * kernels for up to 16 calls, without waiting for them, and counts the
 * (unit, patch) evaluations that its separable phase model declined
 * (evaluated pixel by pixel).  out (host [3], may be NULL) receives the
 * recorded calls' mean times in milliseconds (moments, evaluation) and the
 * count over them, after waiting for their events; then timing stays on or
 * goes off by enable.  No counterpart in the reference, which times its
 * kernels from Python with CuPy events (hardware.py:1944-1980).
 */
int mosaic_kin_timing(int enable, float* out);

/*
 * Adjoint of mosaic_kin_sum with respect to the atoms' sample-frame
 * positions: given grad_field = dL/dRe(A) + i dL/dIm(A) (device
 * (n_frames, Nz, Ny) complex64) for a real loss L of the field A that
 * mosaic_kin_sum produced from the same workspace, adds dL/du (device (N, 3)
 * float32, per Angstrom, original atom order) into grad_pos.  Phases are
 * recomputed, not stored.  No counterpart in the reference, which has no
 * gradient path; it serves displacement-field fitting.  Three forms, chosen
 * like the forward's (`path` 0 automatic, 2 row recurrence, 3 matrix-core,
 * 4 moment expansion): the moment form contracts each atom's expansion
 * coefficients with per-patch pixel moments of the weights; the matrix-core
 * form contracts per-patch weight matrices with the atoms' separable
 * phasors on the FP16-split MFMA; the row form walks row blocks on the VALU.
 * Declines (MOSAIC_ERR_UNSUPPORTED) outside the far-field bounds of the
 * chosen form.  With `scratch` (device, at least mosaic_kin_adjoint_scratch
 * bytes; NULL or smaller: none) the moment form shares each unit's pixel
 * moments among its 1, 2 or 4 tiles, as the forward's moment sum does.
 */
int mosaic_kin_adjoint(int64_t atom_count, int n_species,
                       const int64_t* species_counts /* host */, double k, int n_frames,
                       const float* rot /* host */, const float* trans /* host */, int Ny,
                       int Nz, const double* stage_stats /* host [2] */,
                       const double* pixel_geom /* host [8] */, const float* grad_field,
                       float* grad_pos, void* workspace, size_t workspace_bytes,
                       void* scratch, size_t scratch_bytes, int path,
                       mosaic_stream_t stream);

/*
 * Bytes of the scratch that lets mosaic_kin_adjoint's moment form share pixel
 * moments across units of tiles: 360 B per tile, frame and 32 x 32 patch.  No
 * counterpart in the reference (an adjoint of this build's own).
 */
int mosaic_kin_adjoint_scratch(int64_t atom_count, int n_species,
                               const int64_t* species_counts /* host */, int n_frames, int Ny,
                               int Nz, size_t* bytes);

/*
 * compute_scattering_cffi on the GPU (Beam.py:629-816): the same arguments
 * with the same meaning, each array a device pointer -- positions (N, 3)
 * float64 lab metres, per-atom WK rows (N, 11) float32, f0(0), f', f'' and the
 * incident amplitude (N,) float32 each, pixel coordinates (Ny*Nz,) float64
 * metres each -- accumulating the sum into the float32 device outputs
 * out_r / out_i (+=), ordered on `stream`.  Exact per pair, like the
 * reference: no far-field expansion, pairs with r == 0 skipped.  The pair
 * geometry and phase are FP64 (the phase as Theta0 + K (x + r - R0) with
 * r - R0 in cancellation-free form, so it is free of the reference's
 * ~1e-5 rad fmod(k (x + r), 2 pi) rounding), the form factor, polarization
 * and phasor FP32, the accumulation FP64.  `workspace` is device memory of
 * at least mosaic_scatter_workspace bytes (slice partials).  Returns
 * MOSAIC_OK, MOSAIC_ERR_ARGS or MOSAIC_ERR_WORKSPACE where the reference
 * returns 0 or 1 (allocation failure, Beam.py:5768-5769).
 */
int mosaic_scatter_workspace(int atom_count, int Ny, int Nz, size_t* bytes);

int mosaic_scatter(int atom_count, const double* positions, const float* f0_params,
                   const float* f0_zero, int remove_forward, const float* s_anom_real,
                   const float* s_anom_imag, const float* initial_amp_r,
                   const float* initial_amp_i, int Ny, int Nz, const double* coords_x,
                   const double* coords_y, const double* coords_z, double k_val,
                   int apply_pol, double pol_deg, double pol_ex, double pol_ey,
                   double pol_ez, int apply_spherical_decay, float* out_r, float* out_i,
                   void* workspace, size_t workspace_bytes, mosaic_stream_t stream);

/*
 * mosaic_scatter without a workspace, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation).  No counterpart in the reference.  The slice
 * partials live in a fixed per-device scratch of 64 MiB inside the library,
 * which each GPU it runs on holds until the process ends; the call allocates
 * nothing.  Detectors of more than 4,194,304 pixels (2048 x 2048) return
 * MOSAIC_ERR_TOO_LARGE; fewer atom slices are used where the scratch holds
 * fewer, so the result equals mosaic_scatter's to FP64 summation order.
 * Calls that overlap on one GPU (different streams or threads) share the
 * scratch and race.
 */
int mosaic_scatter_noalloc(int atom_count, const double* positions, const float* f0_params,
                           const float* f0_zero, int remove_forward, const float* s_anom_real,
                           const float* s_anom_imag, const float* initial_amp_r,
                           const float* initial_amp_i, int Ny, int Nz, const double* coords_x,
                           const double* coords_y, const double* coords_z, double k_val,
                           int apply_pol, double pol_deg, double pol_ex, double pol_ey,
                           double pol_ez, int apply_spherical_decay, float* out_r,
                           float* out_i, mosaic_stream_t stream);

/*
 * Incident amplitude per atom by bilinear sampling of the beam's entrance
 * field E0 (NyB x NzB complex64, 'ij' layout) at the lab position
 * rot @ pos + trans; zero outside the grid.  The kinematic case of
 * ein_bilinear_kernel / _ein_bilinear_cpu with tau = phi = 0
 * (Beam.py:956-1045, 2439-2578).
 */
int mosaic_ein_bilinear(int64_t atom_count, const float* pos, const float* rot /* host */,
                        const float* trans /* host */, const float* E0 /* device */,
                        int NyB, int NzB, double du, double dv, double uc,
                        double vc, float* amp_out, mosaic_stream_t stream);

/*
 * Depth-dependent entrance field, for a +x beam on the beam grid (NyB x NzB,
 * pitches du, dv [A], centre indices uc, vc).  Replaces the reference's
 * precompute_depth_ein_all_chunks (Beam.py:4963-5238) with its beam-column map
 * (_compute_beam_column_A_map_cpu/_gpu, Beam.py:7118-7282), depth window
 * (_compute_global_depth_bounds, Beam.py:5583-5620) and E0 * tau * phi
 * sampler (ein_bilinear_kernel / _ein_bilinear_cpu, Beam.py:956-1045,
 * 2439-2578).  Positions are float32 sample-frame Angstrom with the pose
 * lab = rot @ pos + trans (host float32); every call works in one caller
 * workspace of mosaic_ein_depth_workspace bytes, and mosaic_ein_depth reads
 * the maps that the last mosaic_ein_column_map on the same workspace made.
 */
int mosaic_ein_depth_workspace(int NyB, int NzB, size_t* bytes);

/* (s_min, s_max) = min and max of the atoms' lab x [A] (host out [2]);
 * (0, 1) when degenerate.  Synchronizes the stream. */
int mosaic_ein_depth_bounds(int64_t atom_count, const float* pos, const float* rot /* host */,
                            const float* trans /* host */, int NyB, int NzB, double* bounds,
                            void* workspace, size_t workspace_bytes, mosaic_stream_t stream);

/* Beam-column map: triangular-shaped-cloud deposit of each in-grid atom's
 * forward factors fwd[2 * species] = (f0(0) + f', f'') (host float32), then
 * phi = -k C sum_re, tau = max(k C sum_im, 0), C = r_e lambda^2 / (2 pi du dv),
 * an optional Gaussian blur of kernel_radius pixels (<= 16; circular, kernel
 * at the grid corner), A = exp(-tau + i phi) (written to A_out, device
 * NyB x NzB complex64, when not NULL) and the sampled maps
 * tau' = -ln(|A| + 1e-20), phi' = arg A kept in the workspace. */
int mosaic_ein_column_map(int64_t atom_count, const float* pos, const uint8_t* species,
                          int n_species, const float* fwd /* host */, const float* rot /* host */,
                          const float* trans /* host */, int NyB, int NzB, double du, double dv,
                          double uc, double vc, double wavelength_A, int kernel_radius,
                          float* A_out, void* workspace, size_t workspace_bytes,
                          mosaic_stream_t stream);

/* Per-atom Ein = E0 exp(-f tau') exp(i f phi') (E0, tau', phi' bilinear at
 * the atom's (u, v); f = clip((x - s_min) / (s_max - s_min), 0, 1)) into
 * amp_out (device (N,) complex64); zero outside the grid. */
int mosaic_ein_depth(int64_t atom_count, const float* pos, const float* rot /* host */,
                     const float* trans /* host */, const float* E0 /* device */, int NyB,
                     int NzB, double du, double dv, double uc, double vc, double s_min,
                     double s_max, float* amp_out, const void* workspace,
                     size_t workspace_bytes, mosaic_stream_t stream);

/* M-channel incident amplitude of the general kernel (Beam.py:2283-2309):
 * amp[j] = sum_m E[j * M + m] e^{2 pi i g_m . r_j} with r_j = rot @ pos_j +
 * trans in metres (pos float32 sample-frame Angstrom, rot/trans host float32,
 * trans in Angstrom), g (host float64, M x 3) in cycles per metre, M <= 64;
 * E (N, M) and amp_out (N,) complex64 device.  The phase is formed and
 * reduced in FP64.  The result feeds any sum as the atoms' amplitudes. */
int mosaic_mbeam_amplitude(int64_t atom_count, const float* pos, const float* rot /* host */,
                           const float* trans /* host */, const float* E, int M,
                           const double* g /* host */, float* amp_out, mosaic_stream_t stream);

/* Hardware sine and cosine of 2 pi x, x in revolutions (v_sin_f32 /
 * v_cos_f32), as the sum kernels use them, on n device floats; for
 * characterising their accuracy over the argument ranges the kernels feed.
 * No counterpart in the reference. */
int mosaic_sincos_rev(int64_t n, const float* x, float* sin_out, float* cos_out,
                      mosaic_stream_t stream);

/* ----------------------------------------------------------- propagation */

/* Thin lens exp(-i k r^2 / 2f) with optional CRL absorption
 * (_apply_thin_lens_box, Optics.py:49-130). */
typedef struct {
    double focal_length_m;  /* f [m] of the whole stack */
    double parab_coef;      /* mu N / (2 R_apex) [1/m^2]; 0 = no parabolic absorption */
    float uniform_amp;      /* exp(-N t / sigma); 1 = no uniform absorption */
    float pad_;
} mosaic_lens_t;

/* Workspace bytes for one angular-spectrum step on an (Nz2, Ny2) padded grid
 * (the padded field, the FFT work area and the pruned transform's tables). */
int mosaic_asp_workspace(int Ny2, int Nz2, size_t* bytes);

/*
 * One angular-spectrum propagation step over z [m] of the (Nz, Ny) complex64
 * device field `field_in` (pitches dy along columns, dz along rows, metres):
 * optional thin lens (NULL = none) applied first, pad to (Nz2, Ny2) (edge
 * replication when pad_edge, else the constant pad_re + i pad_im), FFT,
 * transfer function, inverse FFT, centre crop into `field_out` (Nz, Ny).
 * Replaces one pass of _angular_spectrum_propagate_gpu/_cpu
 * (Beam.py:7829-8022) with prop_mul_kernel / prop_mul_cpu
 * (Beam.py:4739-4900); the pad size and sub-stepping rules stay with the
 * caller (_choose_optimal_pad / _alias_free_step, Beam.py:4632-4736).
 * `method`: 0 automatic (the pruned transform where Ny2 and Nz2 are powers
 * of two and Ny, Nz <= 8192, else the padded hipFFT round trip), 1 padded
 * hipFFT round trip, 2 pruned transform (MOSAIC_ERR_UNSUPPORTED where its
 * size rules fail).  The pruned transform never forms the padded grid: it
 * transforms the data rows, then each frequency column (transfer function
 * and inverse onto the crop rows fused), then the crop rows.
 * field_in and field_out may alias.
 */
int mosaic_asp_propagate(const float* field_in, int Ny, int Nz, double dy, double dz,
                         double wavelength, double z, int Ny2, int Nz2, int pad_edge,
                         float pad_re, float pad_im, const mosaic_lens_t* lens /* host */,
                         int method, float* field_out, void* workspace, size_t workspace_bytes,
                         mosaic_stream_t stream);

/* Thin lens in place on the (Nz, Ny) device field, pitches dx (columns), dy
 * (rows) in metres (_apply_thin_lens_box, Optics.py:49-130). */
int mosaic_lens_apply(float* field, int Ny, int Nz, double dx, double dy, double wavelength,
                      const mosaic_lens_t* lens /* host */, mosaic_stream_t stream);

/* -------------------------------------------- optics elements, detector */

/* Two-bounce Bragg magnifier, a geometric resampling (_apply_bragg_magnifier_2b,
 * Optics.py:132-209): output pixel (x, y) samples the input at
 * (cx + (x - cx) / Mx, cy + (y - cy) / My), nearest (order 0) or bilinear,
 * zeros outside or edge-clamped (pad_edge), times sqrt(reflectivity)
 * [/ sqrt |Mx My| when conserve_energy] e^{i phase_shift}.  (Ny, Nx)
 * complex64 device fields; field_in and field_out must differ. */
int mosaic_bragg_magnifier(const float* field_in, float* field_out, int Ny, int Nx, double Mx,
                           double My, double reflectivity, double phase_shift, int order,
                           int pad_edge, int conserve_energy, mosaic_stream_t stream);

/* Angular filter on the field's 2-D spectrum (_apply_angular_filter_kspace,
 * Optics.py:211-336): angles theta = k_perp / k from fftfreq, shifted by the
 * centre, rolled by roll_deg; mask 1 / (1 + r^2) (Darwin), r <= 1 (top-hat)
 * or 1 / sqrt(1 + r^(2 order)) (Butterworth), r the normalised deviation of
 * the 1-D slit, the ellipse or the circle; then sqrt(transmission)
 * e^{i phase_shift}.  In place on the (Ny, Nx) complex64 device field, pitches
 * dx (columns), dy (rows) [m]; workspace: the FFT's work area. */
typedef struct {
    double center_x_mrad, center_y_mrad;
    double half_angle_x_mrad, half_angle_y_mrad;
    double transmission, phase_shift, roll_deg;
    int mode_1d;       /* 1: slit along the rolled pass axis */
    int elliptical;    /* 2-D shape: 1 elliptical, 0 circular */
    int rolloff;       /* 0 top-hat, 1 Darwin, 2 Butterworth */
    int order;         /* Butterworth order */
} mosaic_angular_filter_t;

int mosaic_angular_filter_workspace(int Ny, int Nx, size_t* bytes);
int mosaic_angular_filter(float* field, int Ny, int Nx, double dx, double dy, double wavelength,
                          const mosaic_angular_filter_t* filter /* host */, void* workspace,
                          size_t workspace_bytes, mosaic_stream_t stream);

/* Fresnel zone plate as a real-space transmission (_apply_fresnel_zone_plate,
 * Optics.py:338-425): zones r_n = sqrt(n lambda f1), f1 = D dr_N / lambda;
 * amplitude (even zones open), phase (odd zones -1) or ideal (thin-lens phase
 * at f1 / order) inside the diameter, minus a central stop, with an optional
 * cosine edge taper; times sqrt(efficiency).  In place. */
typedef struct {
    double outermost_zone_width_nm, diameter_um, efficiency;
    double central_stop_fraction, edge_taper;
    int order;
    int type;          /* 0 amplitude, 1 phase, 2 ideal */
} mosaic_zone_plate_t;

int mosaic_zone_plate(float* field, int Ny, int Nx, double dx, double dy, double wavelength,
                      const mosaic_zone_plate_t* zp /* host */, mosaic_stream_t stream);

/* Aperture of width (or diameter) width_mm centred on the field, square or
 * circular (_apply_aperture, Optics.py:427-471).  In place. */
int mosaic_aperture(float* field, int Ny, int Nx, double dx, double dy, double width_mm,
                    int circular, mosaic_stream_t stream);

/* Detector point-spread function (apply_psf, Detector.py:873-1087): the
 * intensity |E|^2 (or the amplitude |E| when `amplitude`) of the (rows, cols)
 * complex64 device field convolved in FP64 with `kernel` (host, kernel_size^2,
 * odd size, normalised by the caller), boundary modes per axis as
 * scipy.ndimage: 0 constant, 1 reflect, 2 nearest, 3 mirror, 4 wrap; clipped
 * at 0; the field becomes sqrt(result) (or result) with its phase kept.  In
 * place; synchronizes the stream. */
int mosaic_psf_workspace(int rows, int cols, int kernel_size, size_t* bytes);
int mosaic_psf_apply(float* field, int rows, int cols, const double* kernel /* host */,
                     int kernel_size, int mode_rows, int mode_cols, int amplitude,
                     void* workspace, size_t workspace_bytes, mosaic_stream_t stream);

/* Defected crystal for the surrogate's training images: base (N, 3) float32
 * crystal-frame positions [A] plus one dislocation (isotropic elasticity),
 * a hydrostatic strain and one stacking fault, rotated into the sample frame,
 * into out (N, 3).  params (host, 32 floats): core[3], line[3], e1[3], e2[3],
 * b_screw, b_edge, nu, strain, fault_normal[3], fault_shift[3], R[9]
 * (row-major crystal -> sample), fault_offset (plane fault_normal . r = offset).  Device-side counterpart of the reference's
 * host defect fields (Defects.py:3951-3990 for the fault rule). */
int mosaic_defect_positions(int64_t n, const float* base, const float* params /* host */,
                            float* out, mosaic_stream_t stream);

/* --------------------------------------------------------------- surrogate */
/* Kernels of the learned rocking-scan surrogate.  None has a counterpart in
 * the reference, which has no learned model; they serve the surrogate's
 * training and inference (mosaic_amd/surrogate.py).  Arrays are float32
 * device memory; `batch` runs independent problems (ensemble members) at the
 * given element strides. */

/* C = alpha op(A) op(B) + beta C, BF16 MFMA products with FP32 accumulation;
 * op(A)[m][k] = A[m*lda+k] (transA = 0) or A[k*lda+m]; op(B)[k][n] =
 * B[k*ldb+n] (transB = 0) or B[n*ldb+k]. */
int mosaic_sg_gemm(int transA, int transB, int M, int N, int K, const float* A, int lda,
                   long long strideA, const float* B, int ldb, long long strideB, float* C,
                   int ldc, long long strideC, int batch, float alpha, float beta,
                   mosaic_stream_t stream);

/* Z += bias[col % bias_period] over (rows, cols); Y = SiLU(Z) (act 1) or Z (act 0). */
int mosaic_sg_bias_act(float* Z, float* Y, const float* bias, long long rows, int cols,
                       int bias_period, long long strideZ, long long strideBias, int batch,
                       int act, mosaic_stream_t stream);

/* dZ = dY * act'(Z) over n elements. */
int mosaic_sg_act_bwd(const float* Z, const float* dY, float* dZ, long long n, int act,
                      mosaic_stream_t stream);

/* dbias[p] = sum of dZ over rows and over the columns c with c % period = p. */
int mosaic_sg_bias_grad(const float* dZ, float* dbias, long long rows, int cols, int period,
                        long long strideZ, long long strideBias, int batch,
                        mosaic_stream_t stream);

/* Pixel shuffle of a 2x2 stride-2 transposed convolution, channels last:
 * G[(n H + h) W + w][(dy 2 + dx) Co + c] <-> Y[n][2h+dy][2w+dx][c]
 * (backward = 0: G -> Y; 1: Y -> G). */
int mosaic_sg_shuffle2(float* G, float* Y, int nimg, int H, int W, int Co, long long stride,
                       int batch, int backward, mosaic_stream_t stream);

/* loss[b] += mean(w (P_b - T)^2); dP_b = 2 w (P_b - T) / n, with w = 1, or
 * w = row_weight[i / row_len] for element i when row_weight is given (device,
 * n / row_len entries; row_len must divide n). */
int mosaic_sg_mse(const float* P, const float* T, float* dP, float* loss, long long n,
                  long long strideP, int batch, const float* row_weight, long long row_len,
                  mosaic_stream_t stream);

/* Adam step with decoupled weight decay (bias-corrected, 1-based step). */
int mosaic_sg_adam(float* w, const float* g, float* m, float* v, long long n, float lr,
                   float beta1, float beta2, float eps, int step, float weight_decay,
                   mosaic_stream_t stream);

/* Per frame: ensemble mean (frames, npix) and the members' spread, the
 * root mean over pixels of their variance (frames). */
int mosaic_sg_ensemble_stats(const float* P, float* mean, float* spread, int members,
                             int frames, int npix, long long strideP, mosaic_stream_t stream);

/* Weights of one ensemble packed for inference: W (members, K, N) float32
 * -> per member [ceil(N/32)][ceil(K/16)][64][8] bf16 (2 B each, K and N
 * zero-padded to multiples of 16 and 32), the matrix cores' operand order;
 * `packed` holds members x that many elements. */
int mosaic_sg_pack(int members, int K, int N, const float* W, long long strideW, void* packed,
                   mosaic_stream_t stream);

/* The surrogate network 11 -> 256 -> 512 -> 8x8x64 -> 16x16x64 -> 32x32x32
 * -> 64x64x16 -> 128x128x1 (SiLU after every layer but the last; each
 * decoder stage a 2x2 stride-2 transposed convolution), packed weights
 * (mosaic_sg_pack of the (K, N) matrices 11x256, 256x512, 512x4096, 64x256,
 * 64x128, 32x64, 16x4) and float32 biases (members, N or Cout). */
typedef struct {
    const void* W1;
    const float* b1;
    const void* W2;
    const float* b2;
    const void* W3;
    const float* b3;
    const void* D1;
    const float* d1;
    const void* D2;
    const float* d2;
    const void* D3;
    const float* d3;
    const void* D4;
    const float* d4;
} mosaic_sg_net;

/* Bytes of device workspace mosaic_sg_infer needs (the second layer's
 * activations in operand order). */
int mosaic_sg_infer_workspace(int members, int batch, size_t* bytes);

/* Forward pass of every member on batch inputs X (batch, 11) in two launches:
 * out (members, batch, 128 * 128) float32, the same function as the
 * layer-by-layer GEMMs (BF16 operands, FP32 accumulation).  `post` (host, 4
 * floats, or NULL for the raw output) maps each output y to
 * post[2] exp(post[0] y + post[1]) + post[3]: the log-intensity target's
 * inverse, fused into the store. */
int mosaic_sg_infer(int members, int batch, const float* X, const mosaic_sg_net* net,
                    const float* post /* host [4] or NULL */, void* workspace,
                    size_t workspace_bytes, float* out, mosaic_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MOSAIC_AMD_H */
