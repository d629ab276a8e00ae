/* nicr.h -- C ABI of the MI355X Ni-Cr phase-field time-step library (libnicr.so).
 *
 * The reference solvers have no plugin or FFI layer: their in-process interface
 * is `class Model` (mpi2d/NiCr2D_grains.cpp:194-711, mpi3d-slab/NiCr3D_grains.cpp:336-978).
 * Every entry point below replaces one Model member or public field; the citation
 * on each declaration names the member it stands in for.
 *
 * Conventions
 *  - Plain C, plain pointers and sizes. No torch or HIP types cross this boundary.
 *  - Host arrays hold INTERIOR cells only, in the reference's storage order without
 *    halos: 3D index ((i-1)*ny + (j-1))*nz + (k-1) (Z fastest, mpi3d-slab:165-167);
 *    2D index (i-1)*ny + (j-1) (Y fastest, mpi2d:149). Values are IEEE binary64.
 *  - Device buffers belong to the context; host pointers are borrowed for one call.
 *  - One context per host thread; calls on one context are serialized; no callbacks.
 *  - Every function returns NICR_OK (0) or a negative NICR_E_* code; the message
 *    is then available from nicr_last_error().
 */
#ifndef NICR_H
#define NICR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NICR_ABI_VERSION 2

/* Return codes. */
#define NICR_OK 0
#define NICR_E_INVALID (-1)   /* bad argument or configuration */
#define NICR_E_DEVICE (-2)    /* HIP runtime failure */
#define NICR_E_STATE (-3)     /* call not valid in the current state */

/* Step status (nicr_step_report.status). Mirrors the three runtime_error texts
 * thrown by Model::step (mpi2d:573, mpi2d:585-588, mpi2d:352-354). */
#define NICR_STEP_OK 0
#define NICR_STEP_NONFINITE 1      /* "Nonfinite trial field ..." */
#define NICR_STEP_OUT_OF_BOUNDS 2  /* "Out-of-bounds trial at (...)" (bounds=stop) */
#define NICR_STEP_ETA_RANGE 3      /* "Eta trial outside [0,1] ..." */

/* Enumerations of the reference's string options (mpi2d:48-50). */
#define NICR_BC_CONCENTRATION 0
#define NICR_BC_MU 1
#define NICR_BC_CLOSED 2
#define NICR_CK_LEGACY 0
#define NICR_CK_OFF 1
#define NICR_BOUNDS_STOP 0
#define NICR_BOUNDS_LEGACY 1
#define NICR_NUC_KSP 0
#define NICR_NUC_CENTER 1
#define NICR_NUC_OFF 2

/* Arithmetic of the device step. FP64 reproduces the reference's binary64
 * operation sequence; MIXED evaluates stencils and chemical potentials in
 * binary32 and stores eta in binary32, while O/Cr/phi state, thresholds, the
 * inlet flux and all budgets stay binary64. */
#define NICR_PRECISION_FP64 0
#define NICR_PRECISION_MIXED 1

/* Field identifiers for upload/download. */
#define NICR_FIELD_O 0        /* Model::o */
#define NICR_FIELD_CR 1       /* Model::cr */
#define NICR_FIELD_PHI 2      /* Model::p */
#define NICR_FIELD_GB 3       /* Model::gb (static mask, or eta-derived after refresh) */
#define NICR_FIELD_MU_O 4     /* Model::muO  (valid after nicr_refresh) */
#define NICR_FIELD_MU_CR 5    /* Model::muCr (valid after nicr_refresh) */
#define NICR_FIELD_M_O 6      /* Model::mo   (valid after nicr_refresh) */
#define NICR_FIELD_M_CR 7     /* Model::mc   (valid after nicr_refresh) */
#define NICR_FIELD_ETA_S2 8   /* Model::eta_s2 (sum of eta^2; eta mode) */
#define NICR_FIELD_ETA_PAIRS 9 /* mpi2d Model::eta_pairs (sum over pairs of eta_a^2 eta_b^2) */
#define NICR_FIELD_ETA0 64    /* NICR_FIELD_ETA0 + g is eta field g (0-based) */

/* Mirror of the reference Config (mpi2d:39-64, mpi3d-slab:43-72) as a POD.
 * `dim` selects whose step semantics apply: 2 = mpi2d, 3 = mpi3d-slab. */
typedef struct nicr_config {
    int32_t abi_version;      /* must be NICR_ABI_VERSION */
    int32_t dim;              /* 2 or 3 */
    int32_t nx, ny, nz;       /* nz is ignored (taken as 1) when dim == 2 */
    int32_t steps, output_every, left_edge_sites;
    uint64_t random_seed;
    double dx, dy, dz, dt;
    double grain_diameter_grid, gb_fwhm_grid;
    double gb_factor_o, gb_factor_cr;
    double radius_grid, exclusion_gap_grid, nucleation_period;
    double mu_res, surface_c_ref;
    int32_t bc, ck, bounds, nucleation;
    int32_t grains, oxygen_diffusion, grain_evolution, frozen_eta_cache;
    double eta_W, eta_kappa, eta_L, eta_oxide_penalty;
    int32_t max_eta_fields;
    int32_t n_eta;            /* eta fields held (0 unless grains && grain_evolution) */
    int32_t precision;        /* NICR_PRECISION_* */
    int32_t reserved0;
    /* Derived eta coefficients (mpi2d:88-103); filled by nicr_config_derive(). */
    double eta_K, eta_A, eta_scale;
    /* Inlet O concentration of the concentration BC: Physics::layerO = 0.0023
     * (mpi2d:69, used at mpi2d:216-221,408-411). Not a reference option: the
     * default is the reference's constant; ensembles vary it per instance as the
     * O supply. */
    double layer_o;
} nicr_config;

/* Budget (mpi2d:193, mpi3d-slab:217). The reference keeps long double sums; the
 * device keeps each as an unevaluated double-double (hi + lo). */
typedef struct nicr_budget {
    double nucO[2], nucCr[2], ckO[2], ckCr[2], clipO[2], clipCr[2], bcO[2], bcCr[2];
    int64_t nuclei, clipped;
} nicr_budget;

/* Outcome of nicr_step. On failure the device state equals the state after the
 * last accepted step, and budgets are those of that step (mpi3d-slab:811-820). */
typedef struct nicr_step_report {
    int32_t status;           /* NICR_STEP_* */
    int32_t steps_done;       /* accepted steps in this call */
    int32_t i, j, k;          /* 1-based first failing cell in the reference's loop order */
    int32_t grain;            /* 1-based eta index for NICR_STEP_ETA_RANGE */
    double o, cr, phi;        /* trial values at the failing cell (OUT_OF_BOUNDS) */
    double eta_value;         /* trial eta value (ETA_RANGE) */
} nicr_step_report;

/* Reduction of one field over the interior (Model::sum, mpi2d:229-232; log()
 * extrema, mpi2d:702-705). sum is a double-double. */
typedef struct nicr_stats {
    double sum[2];
    double min, max;
} nicr_stats;

/* ---- configuration ---- */
/* Fill a config with the reference defaults (Config{} at mpi2d:39-64 for dim 2,
 * mpi3d-slab:43-72 for dim 3). */
int nicr_config_default(nicr_config* cfg, int dim);
/* Compute eta_K, eta_A, eta_scale (eta_K, eta_A, eta_transport_scale: mpi2d:88-103). */
int nicr_config_derive(nicr_config* cfg);

/* ---- lifecycle: replaces Model::Model's storage (mpi2d:209-228, mpi3d-slab:356-381) ---- */
int nicr_create(const nicr_config* cfg, void** ctx);
int nicr_destroy(void* ctx);
const char* nicr_last_error(void);
/* Device bytes held by the context. */
int nicr_device_bytes(void* ctx, size_t* bytes);

/* ---- fields: replaces direct access to Model::o/cr/p/gb/eta (public Field members) ---- */
int nicr_upload_field(void* ctx, int field, const double* host);
int nicr_download_field(void* ctx, int field, double* host);
/* Replaces assignment to Model::b and Model::initialO/Cr. */
int nicr_set_budget(void* ctx, const nicr_budget* b);
int nicr_get_budget(void* ctx, nicr_budget* b);

/* ---- time stepping ---- */
/* Model::refresh (mpi2d:412-434, mpi3d-slab:597-620): eta mask, mu and M arrays. */
int nicr_refresh(void* ctx);
/* nsteps x Model::step (mpi2d:551-597, mpi3d-slab:761-826) with all-or-nothing
 * acceptance. Stops at the first rejected step. */
int nicr_step(void* ctx, int nsteps, nicr_step_report* report);

/* ---- oxide insertion ---- */
/* Model::nucleate (mpi2d:543-550, mpi3d-slab:752-759): ordered Ksp scan. */
int nicr_nucleate(void* ctx, int* inserted_cells);
/* nicr_nucleate without device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The scan's
 * scratch (per-tile counts, per-plane oxide counts, the candidate list, the
 * accepted centres) lives in one static 32 MiB device buffer per GPU, so the
 * mesh must satisfy tiles + nx + ncell + 4160 <= 8 Mi ints (about 8.3 M cells);
 * larger meshes are refused with NICR_E_INVALID (use nicr_nucleate). Calls that
 * overlap on one GPU share that buffer and race; results are those of
 * nicr_nucleate bit for bit. */
int nicr_nucleate_noalloc(void* ctx, int* inserted_cells);
/* Model::stamp (mpi2d:503-516, mpi3d-slab:704-724); coordinates are 1-based. */
int nicr_stamp(void* ctx, int x, int y, int z, int* changed);
/* Model::initialize_center_oxide (mpi2d:517-542, mpi3d-slab:725-751): one hard
 * oxide disk/sphere of radius cfg.radius_grid about the box centre, non-periodic,
 * on the current state; inserted O/Cr go to nucO/nucCr and nuclei becomes 1.
 * Errors with the reference's messages: radius >= (min side - 1)/2, a nucleus
 * already recorded, or no cell centre covered (the state is then unchanged). */
int nicr_center_seed(void* ctx, int* cells);

/* ---- diagnostics ---- */
/* Model::sum plus extrema of one interior field. */
int nicr_field_stats(void* ctx, int field, nicr_stats* out);
/* Sum over the interior of h(phi) (mpi3d-slab:970). */
int nicr_sum_h_phi(void* ctx, double out_sum[2]);

/* ---- timing support ---- */
/* Mean device duration (ms) of each step kernel over the launches since the last
 * reset, measured with HIP events on the context's stream. names[] gets up to
 * max entries "kernel=ms". Returns the number of entries. */
int nicr_kernel_times(void* ctx, char* buf, size_t buflen);
int nicr_reset_kernel_times(void* ctx);
int nicr_set_timing(void* ctx, int enabled);
/* Block until the device has finished all work queued on the context. */
int nicr_sync(void* ctx);
/* Plain device-to-device copy of `bytes` bytes repeated `reps` times (kernel
 * k_copy_probe); mean ms per copy. The bandwidth reference against which the
 * step kernels' HBM byte counters and roofline fractions are calibrated. No
 * counterpart in the reference. */
int nicr_copy_probe(void* ctx, size_t bytes, int reps, double* ms);

/* Per-cell parts of the 3D grain rows (Model::grain_log and grain_energy,
 * mpi3d-slab/NiCr3D_grains.cpp:546-590) from the accepted device state.
 * energy_terms (ncell): each cell's discrete grain-energy density times the
 * cell volume, binary64 in the host code's operation order, so a long double
 * sum in cell order is the reference's grain energy. extrema: min and max of
 * every eta value, min and max of sum eta^2, max of the GB mask. counts
 * (n_eta + 1): [0] metal cells (phi < 0.5) with no dominant grain, [g] cells
 * whose dominant grain is g (largest eta above 1e-12, first on ties, phi <
 * 0.5). 3D contexts with eta fields only (at most 256). */
int nicr_grain_terms(void* ctx, double* energy_terms, double extrema[5], int64_t* counts);

/* ---- initial 3D grain geometry on the device ----
 * The per-cell distances of Model::voronoi and initialize_eta in 3D
 * (mpi3d-slab/NiCr3D_grains.cpp:395-406,439-467,469-494), which replace the
 * host's O(cells x seeds x 9) loops in the drop-in program's initial state.
 * For every interior cell (1-based i, j, k; centre (i-.5, j-.5, k-.5); order
 * ((i-1)*ny + (j-1))*nz + (k-1)):
 *   owner  the nearest seed under the periodic Y/Z minimum image (0-based,
 *          the first on ties);
 *   db     the GB-mask distance: over the other seeds and their 9 Y/Z images,
 *          min of max(0, (r_image - r_best) / (2 sep));
 *   de[g]  with de != NULL: the same minimum for seed g with r_best from the
 *          owner image's offsets (the eta weights' distance); 0 for the owner.
 * Binary64 in the host code's operation order, so every value is bitwise the
 * host's; exp() of the mask and of the weights stays with the caller (glibc
 * libm defines them). seeds: ng (x, y, z) in grid units. owner, db: ncell
 * entries; de: ng pointers to ncell doubles, or NULL. Uses the current device
 * and at most about 2 GiB of device memory. No counterpart call in the
 * reference. */
int nicr_grain_distances(int nx, int ny, int nz, int ng, const double* seeds, int32_t* owner, double* db,
                         double* const* de);

/* ---- ensembles of 2D instances ----
 * No counterpart in the reference, whose parameter sweeps are separate runs of
 * the 2D program (one Model per process, mpi2d:1148-1236). An ensemble holds n
 * independent 2D instances that share one nicr_config (physics, dt, BCs,
 * binary64) but differ in mesh size and inlet O supply; every instance steps
 * exactly as its own nicr context would (bitwise), with its own accepted state,
 * rejection and budgets. All instances advance together: one persistent kernel
 * over the tiles of all instances plus a per-instance acceptance kernel per
 * step, replayed from a captured HIP graph. Evolving grains (eta) and Ksp
 * insertion are not part of ensembles; the fields are O, CR, PHI and GB. */
typedef struct nicr_batch_instance {
    int32_t nx, ny;           /* instance mesh */
    double layer_o;           /* inlet O of the concentration BC (reference: 0.0023) */
} nicr_batch_instance;

int nicr_batch_create(const nicr_config* common, const nicr_batch_instance* instances, int n, void** batch);
int nicr_batch_destroy(void* batch);
/* Interior cells summed over the instances; device bytes held. */
int nicr_batch_info(void* batch, long long* cells, size_t* device_bytes);
/* Fields of instance `inst` (0-based), interior cells in Model order (x outer). */
int nicr_batch_upload_field(void* batch, int inst, int field, const double* host);
int nicr_batch_download_field(void* batch, int inst, int field, double* host);
int nicr_batch_get_budget(void* batch, int inst, nicr_budget* out);
int nicr_batch_set_budget(void* batch, int inst, const nicr_budget* in);
/* Up to nsteps steps of every instance; an instance stops at its first rejected
 * step with the semantics of nicr_step. reports: n entries, or NULL. */
int nicr_batch_step(void* batch, int nsteps, nicr_step_report* reports);
/* Tiling: tile height tx (1, 2, 4 or 8 rows of 64 cells), persistent grid size
 * (0: resident blocks), graph replay on (1) or plain launches (0). */
int nicr_batch_set_tiling(void* batch, int tx, int grid, int use_graph);
/* Time every tiling candidate on `steps` dry steps (the instances do not
 * change) and keep the fastest; returns it and its ms per ensemble step. */
int nicr_batch_autotune(void* batch, int steps, int* tx, int* grid, double* ms_per_step);

#ifdef __cplusplus
}
#endif
#endif /* NICR_H */
