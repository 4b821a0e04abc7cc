/*
 * bemcs_hip.h -- C ABI of the MI355X (gfx950) backend for the PY-BEMCS
 * DigitalTwinSimulator time step.
 *
 * Every entry point takes plain device pointers, sizes and an explicit
 * hipStream_t (as void*), returns 0 on success or a negative BEMCS_E* code,
 * never throws, and records a message readable with bemcs_last_error()
 * (thread-local).  The library keeps no global mutable state: calls on
 * distinct buffers/handles are re-entrant.  Particle and mesh buffers are
 * owned by the caller; only Poisson solver handles own device memory.
 *
 * Units follow the reference: positions in mm, velocities in m/s, fields in
 * V/m, potentials in V, times in s.
 */
#ifndef BEMCS_HIP_H
#define BEMCS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEMCS_ABI_VERSION 8

#define BEMCS_OK 0
#define BEMCS_EINVAL -1     /* invalid argument (sizes, null pointers) */
#define BEMCS_EHIP -2       /* HIP runtime error */
#define BEMCS_ENOCONV -3    /* Poisson solve did not reach the tolerance */
#define BEMCS_EOVERFLOW -4  /* a particle buffer ran out of capacity */

typedef void* bemcs_stream_t; /* hipStream_t */

int bemcs_abi_version(void);
const char* bemcs_last_error(void);

/* ------------------------------------------------------------------------
 * Drop-in kernels.
 * ---------------------------------------------------------------------- */

/* Replaces push_particles_boris_cpu / push_particles_boris_taichi
 * (Python/physics_engine.py:188-254, :63-144): bilinear gather of E (and B
 * when Bx/By/Bz are non-null; null means a zero field, as the reference
 * never writes B, :432-434), Boris rotation, drift x += v*dt*1000.  Applied
 * n_rep times in one launch, equal to n_rep successive reference calls with
 * the same dt.  The gather and kick use the reference NumPy path's dtype
 * promotion (float32 state, float64 arithmetic), so results are bit-identical
 * to it on the same inputs. */
int bemcs_push_boris(float* x, float* y, float* vx, float* vy, float* vz, int64_t n,
                     const float* Ex, const float* Ey, const float* Bx, const float* By,
                     const float* Bz, int32_t nx, int32_t ny, double dx, double dy,
                     double dt, double q_m, int32_t n_rep, bemcs_stream_t stream);

/* NGP cell counts (Python/physics_engine.py:178-185, :38-60): counts[iy*nx+ix]
 * += 1 with ix = clip(rint(x/f32(dx)), 1, nx-2), likewise iy.  counts is not
 * cleared. */
int bemcs_deposit_ngp_counts(const float* x, const float* y, int64_t n, double dx,
                             double dy, int32_t nx, int32_t ny, uint32_t* counts,
                             bemcs_stream_t stream);

/* rho[c] = f32(f64(rho[c]) + charge_density) applied counts[c] times, which
 * is bit-identical to np.add.at(rho, idx, charge_density) with a Python-float
 * charge density (:185): NumPy evaluates each indexed add that way, and every
 * added value of one deposit is the same. */
int bemcs_counts_to_rho(const uint32_t* counts, int64_t n_nodes, double charge_density,
                        float* rho, bemcs_stream_t stream);

/* Bilinear (cloud-in-cell) deposit of the model manual
 * (Python/docs/PY-BEMCS_Manual.tex:215-230), the deposit the reference's
 * documentation describes beside the NGP deposit its code uses (:178-185).
 * Adds the four corner weights of the first min(n, *n_dev) particles (n_dev
 * may be null) to weights[] as signed 64-bit fixed point, 2^32 units per unit
 * weight.  Index and fractions are the field gather's: g = x/f32(dx),
 * i = clip(floor(g), 0, nx-2), h = f64(g) - i, weights (1-hx)(1-hy),
 * hx(1-hy), (1-hx)hy, hx hy at [iy][ix], [iy][ix+1], [iy+1][ix],
 * [iy+1][ix+1], each rounded to the nearest unit.  Particles with a
 * non-finite coordinate or g outside [-1, n] deposit nothing.  Totals are
 * exact integer sums, independent of particle order.  weights is not
 * cleared. */
int bemcs_deposit_cic(const float* x, const float* y, int64_t n, const int32_t* n_dev, double dx,
                      double dy, int32_t nx, int32_t ny, uint64_t* weights, bemcs_stream_t stream);

/* rho[c] = f32(f64(rho[c]) + charge_density * W[c] 2^-32) for the signed
 * fixed-point weight sums of bemcs_deposit_cic. */
int bemcs_weights_to_rho(const uint64_t* weights, int64_t n_nodes, double charge_density,
                         float* rho, bemcs_stream_t stream);

/* compute_particle_substeps (:928-966): folds max(ds) of the first *n_dev
 * particles into *ds_max_bits (the bit pattern of a non-negative double; zero
 * it first) and writes n_sub = max(1, ceil(ds_max / ds_lim)) to *n_sub_out,
 * ds_lim = frac*min(dx,dy)*1e-3.  dt2 is the caller's dt**2 (Python's pow,
 * which is not always dt*dt). */
int bemcs_substeps(const float* x, const float* y, const float* vx, const float* vy,
                   const float* vz, const int32_t* n_dev, int64_t n_bound,
                   const float* Exy, int32_t nx, int32_t ny, double dx, double dy,
                   double q_m, double dt, double dt2, double frac, uint64_t* ds_max_bits,
                   int32_t* n_sub_out, bemcs_stream_t stream);

/* Poisson(lam) sample of stream (seed, purpose) at `step` -- host function,
 * the injected-particle count of Python/physics_engine.py:1015-1018. */
int64_t bemcs_poisson_count(uint32_t seed, uint32_t purpose, uint32_t step, double lam);

/* Deterministic math, host build (for checks). */
double bemcs_det_exp(double x);
double bemcs_det_log(double x);
void bemcs_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Replaces thermal_conduction_taichi(T, T_new, mask, nx, ny, Fo_x, Fo_y)
 * (Python/physics_engine.py:147-171, called 10 times with T/T_new swapped at
 * :1586-1597): one explicit conduction sweep over device float32 [ny][nx]
 * maps with the reference NumPy loop's semantics (:1598-1613): T_new = T on
 * every node except interior electrode nodes (mask == 1, 1 <= ix <= nx-2,
 * 1 <= iy <= ny-2), which take max(T + dT, 300) with non-electrode
 * neighbours replaced by the node itself, in float32 and the loop's
 * operation order (bit-identical).  The Taichi kernel also updates electrode
 * nodes on the outer ring; the NumPy loop, the parity target, does not.
 * mask is int32 (isBound.astype(np.int32), as the reference passes it). */
int bemcs_thermal_conduction(const float* T, float* T_new, const int32_t* mask, int32_t nx,
                             int32_t ny, float fo_x, float fo_y, bemcs_stream_t stream);

/* Cell-index division.  The push divides positions by float32(dx) and
 * float32(dy) (IEEE, as NumPy).  For a given divisor the library uses a
 * reciprocal-multiply-plus-remainder form instead, after checking on the
 * device that it equals IEEE division for all 2^32 float32 dividends; the
 * first use of a new divisor runs that check (about 2 ms) and synchronises the
 * stream.  BEMCS_FASTDIV=0 in the environment keeps IEEE division.  No
 * reference counterpart (diagnostics of this library).
 *   bemcs_fastdiv_check: number of dividends where the fast form with
 *     reciprocal r differs from IEEE a / b (synchronous).
 *   bemcs_fastdiv_enabled: 1 if the push uses the fast form for b. */
int bemcs_fastdiv_check(float b, float r, uint64_t* mismatches, bemcs_stream_t stream);
int bemcs_fastdiv_enabled(float b, bemcs_stream_t stream);

/* ------------------------------------------------------------------------
 * Device-resident time step.
 * ---------------------------------------------------------------------- */

/* One species, structure of arrays, capacity `cap`.  tag is the ion's
 * persistent id with bit 31 = isCEX (Python/physics_engine.py:413); it is
 * null for electrons.  count points to a device int32. */
typedef struct {
  float* x;
  float* y;
  float* vx;
  float* vy;
  float* vz;
  uint32_t* tag;
  int32_t* count;
  int64_t cap;
} bemcs_species_t;

/* Per-step ion diagnostics accumulated on the device (Python/physics_engine.py
 * :1316-1497). */
typedef struct {
  uint32_t entered;        /* entered_optics_step */
  uint32_t crossed;        /* exit_count_step (also the theta95 sample size) */
  uint32_t exited;         /* any(exited_mask) test */
  uint32_t lost_to_grid;   /* hit & ~OOB */
  uint32_t see_spawned;    /* SEE electrons created */
  uint32_t see_kept;       /* ... that survived the electron purge */
  uint32_t mcc_events;     /* CEX collisions */
  uint32_t crossed_primary;
  uint32_t overflow;       /* bit flags: 1 SEE append, 2 survivor append beyond
                              capacity, 16 count above the launch bound,
                              32 impact heat or damage out of the fixed-point
                              range (non-finite or beyond 2^53 units) */
  uint32_t surface_hits;   /* hit & x > 0.25 mm: ions that heat, sputter or
                              emit SEE (:1359, :1391, :1419) */
  double sum_vx, sum_vx2, sum_v, sum_v2, sum_E;
  double sum_v_primary, sum_v_cex;
} bemcs_ion_diag_t;

/* Fused ion step: n_sub pushes, periodic wrap, first-hit test along the
 * 8-sample segment, OOB, entered/crossed diagnostics, SEE spawning (with the
 * electron purge applied to the spawned electron), removal, CEX collisions,
 * NGP counts and PPC counts of the survivors.  (:1211-1246, :1289-1529,
 * :1682-1758, :1880-1901) */
typedef struct {
  const float* Exy;          /* interleaved (Ex, Ey), [ny][nx][2] */
  const uint8_t* is_bound;   /* [ny][nx] */
  const int32_t* n_sub;      /* device scalar */
  uint32_t* cnt_next;        /* NGP counts of survivors [ny*nx] */
  uint32_t* ppc;             /* PPC counts (trunc index) [ny*nx] */
  float* crossed_angle;      /* |atan2(vy,vx)| in degrees, capacity = ion cap */
  bemcs_ion_diag_t* diag;
  uint32_t* e_cnt_next;      /* electron NGP counts (SEE electrons) */
  double dt;                 /* full step; the substep is dt / n_sub */
  double q_m;
  double Lx, Ly;             /* mm */
  double m_ion, q;
  double see_vth;            /* sqrt(2 q T_see / m_e) */
  double mcc_n0;             /* params n0 (neutral density scale) */
  double mcc_n0a;            /* n0 * a_corr (Roy model) */
  double mcc_rT;             /* Ly * 1e-3 */
  double mcc_vth;            /* sqrt(2 kB Tn / m_ion) */
  float dx32, dy32;          /* float32(dx), float32(dy) in mm */
  float x_entry32, x_exit32, x_plume32;
  float Lx32, Ly32;
  float dt32;                /* float32(dt) for the SEE back-off */
  int32_t nx, ny, periodic_y;
  int32_t neutral_model;     /* 0 = Roy plume, 1 = uniform */
  int32_t cx_model;          /* 0 = reference fit, 1 = log-linear */
  int32_t mcc_enabled;
  uint32_t seed, step;
  /* Electrode-surface effects of impacts (:1359-1449), for every ion that
   * hits an electrode with x > 0.25 mm, at its first impact node (the first
   * electrode node along the 8-sample segment, else the final node), with
   * v2 = f32(f32(vx*vx + vy*vy) + vz*vz):
   *   heat_acc (null unless sim_mode is Thermal/Both) += rint(dT * 2^20),
   *     dT = f32(f32(f32(f32(half_m32 * v2) * w32) / ccell32) * accel32) [K];
   *   damage_acc (null unless Erosion/Both) += rint(damage * 2^24) when
   *     E = f32(f32(half_m32 * v2) / q32) > eth32,
   *     damage = f32(f32(ycoef32 * f32(E - eth32)) * w32).
   * The accumulators are 64-bit fixed point, so their sums do not depend on
   * the order of the ions; bemcs_thermal and bemcs_erosion consume them.
   * see_defer = 1 appends every spawned SEE electron, unpurged, to see_stage
   * (bemcs_see_commit purges and appends them after bemcs_erosion, so the
   * purge sees cells eroded in the same step, as :1417-1449 and :1534-1561
   * order it); 0 purges them here and appends them to e_out. */
  uint64_t* heat_acc;
  uint64_t* damage_acc;
  bemcs_species_t see_stage;
  float half_m32;            /* float32(0.5 * m_ion) */
  float w32;                 /* float32(macro_weight) */
  float ccell32;             /* float32(C_cell) */
  float accel32;             /* float32(thermal_accel) */
  float q32;                 /* float32(q) */
  float eth32;               /* float32(sputter E_th) */
  float ycoef32;             /* float32(sputter Y_coeff) */
  int32_t see_defer;
  /* User tables (lookup_user_cs, :475-484, used at :1398 for the SEE yield
   * and :1714 for the CEX cross-section): cubic B-splines of log10(value) in
   * log10(E / eV), knots t[n] and coefficients c as SciPy's UnivariateSpline
   * holds them, evaluated as FITPACK's splev with log10(E) clipped to
   * [lo, hi] and 10**y = exp(y ln 10) from the deterministic exp.  A null t
   * selects the built-in yield 0.05 + 1e-4 E or cross-section. */
  const double* cx_t;
  const double* cx_c;
  const double* see_t;
  const double* see_c;
  double cx_lo, cx_hi, see_lo, see_hi;
  int32_t cx_n, see_n;
  /* LDS window of the NGP and PPC count histograms (a work-group counts its
   * particles' nodes in a window of W x H nodes anchored at their smallest
   * (ix, iy) and adds each touched node to the global counts once; nodes
   * outside the window go straight to global atomics): 0 = 64 x 32, else
   * log2(W) | log2(H) << 8 | 1 << 16 with W x H <= 2048.  Counts are
   * integers, so the window changes no result, only the number of global
   * atomics. */
  int32_t count_window;
  int32_t pad_;
} bemcs_ion_step_t;

int bemcs_ion_step(const bemcs_species_t* in, const bemcs_species_t* out,
                   const bemcs_species_t* e_out, const bemcs_ion_step_t* p,
                   int64_t n_bound, bemcs_stream_t stream);

/* Fused electron step: n_sub pushes, periodic wrap, purge (nearest-node hit
 * or outside [0,Lx]x[0,Ly] or NaN, :1534-1561), removal, NGP counts.
 * With sub_chunks = K > 1 the substeps are split into K chunks, substeps
 * [floor(n_sub j / K), floor(n_sub (j+1) / K)) in chunk j; the fused step then
 * runs only the substeps from chunk sub_first on (earlier chunks ran through
 * bemcs_electron_push_chunk, with cell sorts between them to keep the
 * gathers of a wavefront in a few cells).  Each electron's substeps depend
 * only on its own state, so the split changes no result. */
typedef struct {
  const float* Exy;
  const uint8_t* is_bound;
  const int32_t* n_sub;
  uint32_t* cnt_next;
  uint32_t* overflow;
  double dt;
  double q_m;
  double Ly;
  float dx32, dy32;
  float Lx32, Ly32;
  int32_t nx, ny, periodic_y;
  int32_t sub_first;         /* first chunk the fused step runs (0: all substeps) */
  int32_t sub_chunks;        /* K (<= 1: no split) */
  int32_t count_window;      /* LDS window of the NGP counts, as in bemcs_ion_step_t */
} bemcs_electron_step_t;

int bemcs_electron_step(const bemcs_species_t* in, const bemcs_species_t* out,
                        const bemcs_electron_step_t* p, int64_t n_bound,
                        bemcs_stream_t stream);

/* Chunk `chunk` of p->sub_chunks of the electron substeps (the push of
 * :1251-1283, n_sub read on the device), in place on the first
 * min(*sp->count, n_bound) electrons: no wrap, purge or removal. */
int bemcs_electron_push_chunk(const bemcs_species_t* sp, const bemcs_electron_step_t* p, int32_t chunk,
                              int64_t n_bound, bemcs_stream_t stream);

/* Injection of `n` particles at the end of `sp` (count is advanced), adding
 * their NGP counts to cnt.  kind: 0 = beam ions (:1020-1064), 1 = neutraliser
 * electrons (:1085-1099), 2 = RF source electrons (:1066-1083), 3 = preloaded
 * ions, 4 = preloaded electrons (uniform over non-electrode nodes).
 *
 * Population hold (kinds 3 and 4 with target > 0; no reference counterpart,
 * a synthetic workload option beside the preload): the launch tops the
 * device count up to `target`, adding need = min(max(target - count, 0), n)
 * particles placed and drawn as the preload does, from the streams
 * (seed, 0x12 ions / 0x13 electrons) with ids *id_next .. *id_next + need - 1
 * (a device counter advanced by need; ion ids are 0x40000000 | id, so they
 * never meet the host-assigned ids below 2^30).  The host needs no count.
 *
 * overflow (may be null): bit 8 is set when the count had to be clamped at
 * the capacity (the host's count bound was wrong). */
typedef struct {
  uint8_t const* is_bound;
  uint32_t* cnt;
  double y_lo[2], y_hi[2];   /* sampling band(s); two_holes uses both */
  double x0;                 /* injection x (mm) */
  double v_drift;            /* vx offset (v_bohm) */
  double v_spread;           /* thermal speed multiplying the normals */
  double Lx, Ly;             /* preload box */
  float dx32, dy32;
  int32_t nx, ny;
  int32_t kind;
  int32_t two_band;          /* 1: choose band by a uniform (two_holes) */
  int32_t band_f32;          /* 1: band limits are float32 values */
  uint32_t seed, step;
  uint32_t pid_base;         /* ions: persistent id of the first particle */
  uint32_t pad_;
  uint32_t* overflow;        /* may be null */
  uint32_t* id_next;         /* population hold: device id counter */
  int64_t target;            /* population hold: count to top up to (0: off) */
} bemcs_inject_t;

int bemcs_inject(const bemcs_species_t* sp, const bemcs_inject_t* p, int64_t n,
                 int64_t n_bound_before, bemcs_stream_t stream);

/* Counting sort of a species by gather cell (iy0*nx + ix0); writes to `out`
 * and sets out->count.  scratch must hold nx*ny + ceil(nx*ny/1024) + 1024 +
 * n_bound uint32. */
int bemcs_sort_by_cell(const bemcs_species_t* in, const bemcs_species_t* out, int64_t n_bound,
                       int32_t nx, int32_t ny, double dx, double dy, uint32_t* scratch,
                       bemcs_stream_t stream);

/* The same counting sort by gather tile: tiles of 2^shift_x x 2^shift_y
 * cells (key (iy0 >> shift_y) * tiles_x + (ix0 >> shift_x)), which keeps a
 * wavefront's gathers in a few cache lines with fewer, longer runs to
 * scatter than a per-cell sort.  Order within a tile is unspecified.
 * 0 <= shift <= 8; scratch as bemcs_sort_by_cell. */
int bemcs_sort_by_tile(const bemcs_species_t* in, const bemcs_species_t* out, int64_t n_bound,
                       int32_t nx, int32_t ny, double dx, double dy, int32_t shift_x, int32_t shift_y,
                       uint32_t* scratch, bemcs_stream_t stream);

/* ------------------------------------------------------------------------
 * Poisson: matrix-free MG-PCG on the reference operator.
 * ---------------------------------------------------------------------- */

typedef struct bemcs_poisson_s* bemcs_poisson_t;

/* Builds the reduced SPD system of build_sparse_matrix (:525-608) from the
 * host isBound map (uint8 [ny][nx], 1 = electrode/Dirichlet) and the geometry
 * mode, plus its multigrid hierarchy; allocates device memory.  The operator
 * arrays (node kinds, coarse stencils, inverse diagonals, the dense coarsest
 * inverse) depend only on the mesh and the map: a handle created for a map
 * and mesh that a live handle was built from shares that handle's arrays
 * (bit-identical to building them again) and allocates only its own vectors,
 * so sweep points of one geometry read one operator. */
int bemcs_poisson_create(const uint8_t* is_bound_host, int32_t nx, int32_t ny,
                         int32_t periodic_y, bemcs_poisson_t* out);
/* The same with the size of the coarsest multigrid level chosen: levels halve
 * until one has at most coarse_max nodes (16..4096), whose operator is
 * inverted densely.  bemcs_poisson_create uses 2048 (fewest launches per
 * V-cycle); 512 suits operators rebuilt every few steps (bemcs_poisson_update
 * inverts the coarsest level again, at cost growing as its size cubed). */
int bemcs_poisson_create_opt(const uint8_t* is_bound_host, int32_t nx, int32_t ny, int32_t periodic_y,
                             int32_t coarse_max, bemcs_poisson_t* out);
int bemcs_poisson_free(bemcs_poisson_t h);
int bemcs_poisson_levels(bemcs_poisson_t h);

/* Coarse tile edge of the level-0 V-cycle launches for the handle's own
 * solves and for the solve sets it joins afterwards: 8 (a work-group owns
 * 16 x 16 fine nodes), 16 (32 x 32: a quarter of the partial sums and less
 * halo traffic, the better choice in large batches) or 0 (by mesh size: 16
 * above 600K nodes).  A set takes 16 if any of its handles does, so a case
 * solved alone and in a batch of handles with the same setting gives the
 * same bits.  Solver tuning without a reference counterpart (the reference
 * factorises with SuperLU, Python/physics_engine.py:608); call it between
 * solves (it rebuilds the handle's own solve set). */
int bemcs_poisson_set_level0_tile(bemcs_poisson_t h, int32_t tile);

/* Rebuilds the operator of handle h for a new electrode map of the same mesh
 * (device uint8 [ny][nx]) in place, on the device: the reference re-runs
 * build_sparse_matrix and factorises again after erosion (:1448) and after a
 * thermal-bow rebuild (:1646).  The multigrid hierarchy and the coarsest-level
 * inverse are recomputed into the handle's existing arrays (bemcs_poisson_create
 * runs the same set-up), so results equal those of a handle created from the
 * new map; the next Picard call starts its solve from the current V.  Waits
 * for the set-up on `stream` to report an invalid map.  A handle that shares
 * its operator arrays with other handles first moves to arrays of its own;
 * solve sets made earlier that hold it then refuse to run (BEMCS_EINVAL)
 * until they are made again. */
int bemcs_poisson_update(bemcs_poisson_t h, const uint8_t* is_bound, bemcs_stream_t stream);

/* Solves the full reference system A V = b (laplacian_lu(b), :900) for
 * device f64 vectors; V holds the initial guess on entry.  Stops when
 * ||r||_2 <= rtol * ||b_reduced||_2 on the eliminated system. */
int bemcs_poisson_solve(bemcs_poisson_t h, const double* b, double* V, double rtol,
                        int32_t max_iter, int32_t* iters, double* relres,
                        bemcs_stream_t stream);

/* recalc_poisson (:850-902): `iterations` Picard sweeps with Boltzmann
 * electrons, rho = ion/electron NGP counts converted as bemcs_counts_to_rho
 * (cd_ion first, then cd_e), SuperLU replaced by MG-PCG, then E = -grad V
 * (np.gradient, periodic fix) written interleaved to Exy (float32).  V is
 * f64 [ny][nx] in/out; V_fixed f64; rho_out (float32) receives rho.  When
 * w_ion is non-null, rho comes instead from the bilinear weight sums w_ion and
 * w_e (bemcs_deposit_cic) converted as bemcs_weights_to_rho on a zero rho,
 * ions first; the count pointers are then ignored. */
typedef struct {
  const uint32_t* cnt_ion;
  const uint32_t* cnt_e;
  const double* V_fixed;
  double* V;
  float* rho_out;
  float* Exy;
  double cd_ion, cd_e;
  double coeff;       /* (dx*1e-3)^2 / eps0 */
  double V_plasma, Te_up, n0, q;
  double omega;
  double dx_m, dy_m;  /* mesh spacing in metres for the gradient */
  double rtol;
  int32_t iterations;
  int32_t max_iter;
  const uint64_t* w_ion;  /* bilinear deposit (null: NGP counts) */
  const uint64_t* w_e;
} bemcs_picard_t;

int bemcs_poisson_picard(bemcs_poisson_t h, const bemcs_picard_t* p, int32_t* iters_total,
                         double* worst_relres, bemcs_stream_t stream);

/* Batched Picard: the Poisson solves of several cases (each a handle from
 * bemcs_poisson_create, meshes of any sizes) advance together, one launch
 * per V-cycle level serving every case (launch index z = case), so a sweep
 * of uneven meshes costs one launch chain per PCG iteration instead of one
 * per case.  params, iters_total and worst_relres are host arrays of n
 * entries; iterations, rtol and max_iter must agree across cases.  Results
 * are bit-identical to bemcs_poisson_picard on each handle.  A handle must not
 * be used by two batches, or by a batch and its own calls, at the same time.
 * No reference counterpart (the reference runs sweep cases in separate
 * processes, Python/benchmarks/benchmark_perveance.py:102-113). */
typedef struct bemcs_poisson_batch_s* bemcs_poisson_batch_t;
int bemcs_poisson_batch_create(const bemcs_poisson_t* handles, int32_t n,
                               bemcs_poisson_batch_t* out);
int bemcs_poisson_batch_free(bemcs_poisson_batch_t b);
int bemcs_poisson_batch_size(bemcs_poisson_batch_t b);
int bemcs_poisson_batch_picard(bemcs_poisson_batch_t b, const bemcs_picard_t* params,
                               int32_t* iters_total, double* worst_relres, bemcs_stream_t stream);

/* E = -grad V only (:871-884). */
int bemcs_efield(const double* V, int32_t nx, int32_t ny, double dx_m, double dy_m,
                 int32_t periodic_y, float* Exy, bemcs_stream_t stream);

/* ------------------------------------------------------------------------
 * Diagnostics.
 * ---------------------------------------------------------------------- */

/* np.percentile(angles[:n], q_percent) with linear interpolation evaluated in
 * float32 as NumPy does for float32 data (bit-identical); n read from the
 * device; result written to *out (double), NaN when n <= min_count.
 * scratch is unused (may be null). */
int bemcs_percentile(const float* angles, const uint32_t* n_dev, int64_t n_bound, double q_percent,
                     uint32_t min_count, uint32_t* scratch, double* out, bemcs_stream_t stream);

/* accum += ppc; stats[0] = #cells with ppc > 0, stats[1] = #cells with
 * 0 < ppc < 3 (:1755-1767). */
int bemcs_ppc_stats(const uint32_t* ppc, int64_t n_nodes, double* accum, uint32_t* stats,
                    bemcs_stream_t stream);

/* Per-step bookkeeping of the device-resident step in one launch each (the
 * reference keeps these as host variables of step(), :969-1784).
 * bemcs_step_begin clears the substep reductions, the output species counts,
 * the node arrays (next deposit counts of both species, PPC map), the ion
 * diagnostics record and the PPC statistics, and stores the input species
 * counts (the particles this step pushes) in nin_out[2].
 * bemcs_step_end stores the species counts and substep counts after the step
 * and the minimum of V over the saddle-point nodes (:1655-1676). */
typedef struct {
  uint64_t* ds_bits;         /* [2] */
  int32_t* count_out_i;      /* output buffers' counts */
  int32_t* count_out_e;
  uint32_t* node_zero[3];    /* [n_nodes] each, may be null */
  int64_t n_nodes;
  uint8_t* diag;             /* bemcs_ion_diag_t record */
  int64_t diag_bytes;
  uint32_t* ppc_stats;       /* [2] */
  const int32_t* count_in_i; /* input buffers' counts */
  const int32_t* count_in_e;
  int32_t* nin_out;          /* [2] */
  int32_t* count_out_x;      /* a third count cleared (SEE staging), may be null */
} bemcs_step_begin_t;

typedef struct {
  const int32_t* count_i;
  const int32_t* count_e;
  const int32_t* n_sub;      /* [2] */
  int32_t* counts_out;       /* [2] */
  int32_t* nsub_out;         /* [2] */
  const double* V;
  const int64_t* saddle_idx;
  int32_t n_saddle;
  int32_t pad_;
  double* minpot_out;        /* NaN when n_saddle == 0 */
} bemcs_step_end_t;

int bemcs_step_begin(const bemcs_step_begin_t* p, bemcs_stream_t stream);
int bemcs_step_end(const bemcs_step_end_t* p, bemcs_stream_t stream);

/* ------------------------------------------------------------------------
 * Electrode-surface physics (sim_mode Thermal / Erosion / Both) and RF.
 * ---------------------------------------------------------------------- */

/* Sputter damage and cell failure of one step (:1419-1449), after
 * bemcs_ion_step: skipped entirely when diag->surface_hits is 0 (the
 * reference tests for failed cells only on steps with electrode hits);
 * otherwise damage_map += damage_acc * 2^-24 and damage_acc = 0 on every
 * node, and each node with damage_map > thresh that is still an electrode
 * fails: eroded_depth += dy, is_bound = 0, damage_map = 0, *n_failed += 1
 * (a running total: the caller rebuilds the Poisson operator when it grows,
 * build_sparse_matrix at :1448). */
typedef struct {
  uint64_t* damage_acc;
  double* damage_map;
  double* eroded_depth;
  uint8_t* is_bound;
  const bemcs_ion_diag_t* diag;
  uint32_t* n_failed;
  int64_t n_nodes;
  double thresh;             /* params Thresh */
  double dy;                 /* mm */
} bemcs_erosion_t;

int bemcs_erosion(const bemcs_erosion_t* p, bemcs_stream_t stream);

/* Purge and append of SEE electrons staged by bemcs_ion_step with see_defer:
 * an electron survives unless its nearest node is an electrode (after this
 * step's erosion) or it lies outside [0,Lx]x[0,Ly] or is NaN (:1534-1561);
 * survivors are appended to e_out with their NGP counts in e_cnt_next and
 * counted in diag->see_kept.  n_bound bounds the staged count. */
int bemcs_see_commit(const bemcs_species_t* stage, const bemcs_species_t* e_out,
                     const uint8_t* is_bound, uint32_t* e_cnt_next, bemcs_ion_diag_t* diag,
                     int32_t nx, int32_t ny, float dx32, float dy32, float Lx32, float Ly32,
                     int64_t n_bound, bemcs_stream_t stream);

/* Thermal update of one step (:1563-1617) in one pass over the map:
 * T = T_in + heat_acc * 2^-20 (this step's impact heat), radiative cooling
 * of electrode nodes T -= f32(cool32 * f32(f32(T^4) - t300_4_32)), `sweeps`
 * explicit conduction sweeps over interior electrode nodes (insulated
 * towards non-electrode neighbours, floor 300 K, float32 in the reference
 * loop's operation order, :1598-1613), the 300 K floor on every electrode
 * node, written to T_out (distinct from T_in); heat_clear (the other
 * accumulator of a ping-pong pair, may be null) is zeroed for the next step.
 * The mean of T_out over the nodes of each grid (bit g of grid_bits) is
 * written to T_grids[g] as float32 for grids with at least one node.
 * workspace: bemcs_thermal_workspace_bytes(nx, ny, n_grids) bytes.
 * sweeps <= 16, n_grids <= 8. */
typedef struct {
  const float* T_in;
  float* T_out;
  const uint64_t* heat_acc;  /* may be null (no impact heat) */
  uint64_t* heat_clear;
  const uint8_t* is_bound;
  const uint8_t* grid_bits;
  void* workspace;
  float* T_grids;            /* [n_grids] */
  float cool32;              /* float32(emissivity*sigma*A_cell*dt*accel/C_cell) */
  float t300_4_32;           /* float32(300.0**4) */
  float fo_x32, fo_y32;      /* float32 Fourier numbers */
  int32_t nx, ny, n_grids, sweeps;
} bemcs_thermal_t;

int64_t bemcs_thermal_workspace_bytes(int32_t nx, int32_t ny, int32_t n_grids);
int bemcs_thermal(const bemcs_thermal_t* p, bemcs_stream_t stream);

/* bemcs_thermal without a workspace and without any allocation: p->workspace
 * is ignored and the per-tile partial sums go to a 1 MiB scratch array the
 * library holds in device memory (one per GPU, from the first call on that
 * device until the process ends).  No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  Meshes with bemcs_thermal_workspace_bytes(nx, ny, n_grids) above
 * 1048576 bytes are rejected with BEMCS_EINVAL (at 8 grids, up to 8192
 * tiles of 64 x 32 nodes, e.g. 4096 x 4096).  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (different streams, no
 * ordering between them) race and must not be issued. */
int bemcs_thermal_noalloc(const bemcs_thermal_t* p, bemcs_stream_t stream);

/* RF modulation of a grid potential (:984-991): on every node of grid rf_grid
 * (bit rf_grid of grid_bits), V_fixed = Vdc + v_rf, Vdc being the potential
 * of the last grid covering the node (grid_V: host array of n_grids values). */
int bemcs_rf_set(double* V_fixed, const uint8_t* grid_bits, int64_t n_nodes, int32_t n_grids,
                 const double* grid_V, int32_t rf_grid, double v_rf, bemcs_stream_t stream);

/* Electrode rasterisation of build_domain (:791-843) on the device: node
 * coordinates as np.linspace(0, L, n), for each grid in order the slab
 * gstart + dxbow <= x <= gend + dxbow (dxbow = delta * eta^2, eta per
 * geometry, when |delta| > 1e-6) minus the apertures |y - yc| <= r - max(0,
 * x - gstart - dxbow) tan(chamfer) (clipped at 0); is_bound = 1 and V_fixed
 * = V on the grid's nodes (later grids overwrite), bit g of grid_bits set;
 * column 0 is Dirichlet at V_plasma.  Float64 in the reference's operation
 * order: the map equals NumPy's. */
typedef struct {
  double gstart, gend;       /* x extent of the grid (mm) */
  double V;                  /* grid potential */
  double r;                  /* aperture radius (mm) */
  double tan_cham;           /* np.tan(np.radians(chamfer)) */
  double delta;              /* thermal-bow deflection (mm) */
} bemcs_grid_geom_t;

typedef struct {
  uint8_t* is_bound;         /* [ny][nx] out */
  double* V_fixed;           /* [ny][nx] out */
  uint8_t* grid_bits;        /* [ny][nx] out */
  bemcs_grid_geom_t grids[8];
  double hole_y[2];          /* aperture centres (mm) */
  double Lx, Ly;             /* mm */
  double screen_r, pitch;    /* mm (web position of the bow profiles) */
  double V_plasma;           /* column-0 Dirichlet value */
  int32_t n_grids, n_holes;
  int32_t nx, ny;
  int32_t geometry;          /* 0 half_hole, 1 one_hole, 2 two_holes */
  int32_t pad_;
} bemcs_electrodes_t;

int bemcs_electrodes(const bemcs_electrodes_t* p, bemcs_stream_t stream);

/* Mean of T over the nodes of each grid (bit g of grid_bits), written to
 * T_grids[g] as float32 for grids with at least one node (others are left
 * unchanged); workspace as bemcs_thermal_workspace_bytes(nx, ny, n_grids). */
int bemcs_grid_means(const float* T, const uint8_t* grid_bits, int32_t nx, int32_t ny, int32_t n_grids,
                     void* workspace, float* T_grids, bemcs_stream_t stream);

/* ------------------------------------------------------------------------
 * Batched step kernels: one launch serves n independent cases (meshes and
 * particle counts may differ); launch index z selects the case.  `cases` is a
 * device array of n entries (caller-owned, e.g. one upload per step of all
 * the step's argument structs).  Each case's results are bit-identical to the
 * single-case call with the same arguments.  No reference counterpart: the
 * reference runs sweep cases in separate processes
 * (Python/benchmarks/benchmark_perveance.py:102-113).
 * ---------------------------------------------------------------------- */
typedef struct {
  bemcs_species_t in, out, e_out;
  bemcs_ion_step_t p;
} bemcs_ion_case_t;

typedef struct {
  bemcs_species_t in, out;
  bemcs_electron_step_t p;
} bemcs_electron_case_t;

typedef struct {
  const float* x;
  const float* y;
  const float* vx;
  const float* vy;
  const float* vz;
  const int32_t* n_dev;
  const float* Exy;
  int32_t nx, ny;
  double dx, dy, q_m, dt, dt2, frac;
  uint64_t* ds_max_bits;
  int32_t* n_sub_out;
} bemcs_substeps_case_t;

typedef struct {
  bemcs_species_t sp;
  bemcs_inject_t p;
  int64_t n;               /* particles to inject (0 allowed) */
  int64_t n_bound_before;  /* host bound of the count before injection */
} bemcs_inject_case_t;

typedef struct {
  bemcs_species_t in, out;
  int64_t n_bound;
  int32_t nx, ny;
  double dx, dy;
  uint32_t* scratch;       /* as bemcs_sort_by_cell */
  int32_t shift_x, shift_y; /* tile of 2^shift_x x 2^shift_y cells (0, 0: one cell) */
} bemcs_sort_case_t;

typedef struct {
  const uint32_t* ppc;
  int64_t n_nodes;
  double* accum;
  uint32_t* stats;
} bemcs_ppc_case_t;

typedef struct {
  const float* angles;
  const uint32_t* n_dev;
  double q_percent;
  uint32_t min_count;
  uint32_t pad_;
  double* out;
} bemcs_percentile_case_t;

/* divisors_host: host array [2n] of float32 (dx, dy) per case, for the
 * division check of bemcs_fastdiv_enabled; max_bound: largest host count
 * bound over the cases (sets the launch size). */
int bemcs_ion_step_batched(const bemcs_ion_case_t* cases, int32_t n, int64_t max_bound,
                           const float* divisors_host, bemcs_stream_t stream);
int bemcs_electron_step_batched(const bemcs_electron_case_t* cases, int32_t n, int64_t max_bound,
                                const float* divisors_host, bemcs_stream_t stream);
int bemcs_substeps_batched(const bemcs_substeps_case_t* cases, int32_t n, int64_t max_bound,
                           bemcs_stream_t stream);
/* max_n: largest injected count over the cases */
int bemcs_inject_batched(const bemcs_inject_case_t* cases, int32_t n, int64_t max_n,
                         bemcs_stream_t stream);
/* max_nodes: largest nx*ny over the cases */
int bemcs_sort_by_cell_batched(const bemcs_sort_case_t* cases, int32_t n, int64_t max_bound,
                               int64_t max_nodes, bemcs_stream_t stream);
int bemcs_step_begin_batched(const bemcs_step_begin_t* cases, int32_t n, int64_t max_nodes,
                             bemcs_stream_t stream);
int bemcs_step_end_batched(const bemcs_step_end_t* cases, int32_t n, bemcs_stream_t stream);
int bemcs_ppc_stats_batched(const bemcs_ppc_case_t* cases, int32_t n, bemcs_stream_t stream);
int bemcs_percentile_batched(const bemcs_percentile_case_t* cases, int32_t n,
                             bemcs_stream_t stream);

/* HIP graphs of a launch sequence, for the batched sweep's per-step kernels.
 * No counterpart in the reference, which runs sweep cases as one process each
 * (Python/benchmarks/benchmark_perveance.py:102-113).  After
 * bemcs_graph_capture_begin, launches on `stream` from the calling thread are
 * recorded, not run, until bemcs_graph_capture_end instantiates them into
 * *exec_out (abort != 0 discards them, after a launch that failed).  Kernel
 * arguments are fixed at capture: a caller that replays a graph with
 * bemcs_graph_launch updates the argument structs in device memory that the
 * captured kernels read.  The graphs hold no memset nodes. */
int bemcs_graph_capture_begin(bemcs_stream_t stream);
int bemcs_graph_capture_end(bemcs_stream_t stream, int32_t abort, void** exec_out);
int bemcs_graph_launch(void* exec, bemcs_stream_t stream);
int bemcs_graph_destroy(void* exec);

#ifdef __cplusplus
}
#endif
#endif /* BEMCS_HIP_H */
