/*
 * antimatter-kernel for MI355X — C ABI of the per-step hot path.
 *
 * Every entry point returns an int status (AM_OK = 0, negative on error; no exceptions cross
 * the ABI) and leaves a message for am_last_error().  Arguments are plain pointers and sizes.
 * Citations name the reference interface each entry point replaces (paths relative to the
 * reference repository root).
 *
 * Objects:
 *   am_ctx   — one GPU device and one HIP stream.  Calls on a context (and on the stores made
 *              from it) must be serialised by the caller, as the reference's solvers are called
 *              from the driver's single thread (src/antimatter-kernel.cc:153-197).
 *   am_store — structure-of-arrays particle store resident in HBM; the GPU counterpart of the
 *              ResourceManager's ChargedAntiparticle agents (src/agents/charged_antiparticle.h:14-79).
 *              Particle i keeps index i for its whole life; annihilated particles stay in the
 *              store, flagged, exactly as IsAnnihilated() agents are skipped by every solver.
 */
#ifndef ANTIMATTER_AMD_AM_API_H_
#define ANTIMATTER_AMD_AM_API_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AM_OK 0
#define AM_ERR_INVALID -1  /* invalid argument (mirrors std::invalid_argument in the reference) */
#define AM_ERR_HIP -2      /* HIP runtime error */
#define AM_ERR_CAPACITY -3 /* size exceeds a capacity */
#define AM_ERR_STATE -4    /* call out of order / missing data */

/* Arithmetic used by a solver's pair kernel. */
#define AM_FP64 0 /* FP64 end to end (parity path) */
#define AM_FP32 1 /* FP32 pair math on hi/lo-split coordinates, FP64 accumulation per tile */

/* ParticleSpecies (src/agents/particle_species.h:9-17). */
#define AM_ANTIPROTON 0
#define AM_POSITRON 1
#define AM_ANTIHYDROGEN 2
#define AM_ELECTRON 3
#define AM_PROTON 4
#define AM_ANTIHELIUM 5

typedef struct am_ctx am_ctx;
typedef struct am_store am_store;

/* PenningTrapField (src/fields/penning_trap.h:27-69): B = b0 z_hat,
 * E = (U0 / 2 d^2)(x, y, -2 z), phi = (U0 / 4 d^2)(2 z^2 - rho^2). */
typedef struct {
  double b0;
  double trap_voltage;
  double d; /* characteristic dimension, d = sqrt(z0^2/2 + rho0^2/4) */
} am_trap_field;

/* EnergyBreakdown (src/analysis/energy_analysis.h:16-23). */
typedef struct {
  double kinetic;
  double trap_potential;
  double energy_potential;
  int64_t num_active_particles;
} am_energy_breakdown;

/* EnergyHistory (src/analysis/energy_analysis.h:26-31). */
typedef struct {
  int32_t has_previous;
  double prev_trap_potential;
  double prev_coulomb_potential;
} am_energy_history;

/* ComputationMethod (src/params/interaction_params.h:11-16), plus all-pairs. */
#define AM_METHOD_STANDARD 0
#define AM_METHOD_BARNES_HUT 1
#define AM_METHOD_FAST_MULTIPOLE 2
#define AM_METHOD_ALLPAIRS 3 /* exact all-pairs: the theta = 0 limit of BH/FMM */

/* EnergyComputationSettings (src/params/interaction_params.h:18-25). */
typedef struct {
  int32_t method;
  double cutoff_radius;
  double barnes_hut_theta;
  double fast_multipole_theta;
  double fast_multipole_order;
} am_energy_settings;

/* ---- errors / version ---------------------------------------------------------------------- */
const char* am_last_error(void);
int am_version(void);

/* ---- field --------------------------------------------------------------------------------- */
/* PenningTrapField::MakeFromGeometry (src/fields/penning_trap.cc:15-20). */
int am_trap_make_from_geometry(double b0, double trap_voltage, double z0, double rho0, am_trap_field* out);
/* CyclotronCooling::CoolingTime (src/forces/cooling.cc:16-22). */
double am_cooling_time(double charge, double mass, double b_mag);
/* GetDefaultSpeciesProperties (src/agents/particle_species.h:27-46). */
int am_species_properties(int species, double* mass, double* charge);

/* ---- initial conditions (host) --------------------------------------------------------------- */
/* PenningTrapField::SetParticlePositions (src/fields/penning_trap.cc:118-161): counts[s] particles
 * of species s in enum order; out: 3 * sum(counts) doubles.  Uses libstdc++ <random>. */
int am_set_particle_positions(const uint64_t counts[6], const double radii[3], int seed, double* out);
/* PenningTrapField::SetParticleVelocities (src/fields/penning_trap.cc:51-115). */
int am_set_particle_velocities(const uint64_t counts[6], const double temperatures[6], int seed, double* out);
/* PhysicsParams dt rule (src/params/physics_params.cc:60-74): listed[s] != 0 for species named
 * in particle_populations. */
double am_derive_dt(const int32_t listed[6], double b0);

/* Seeded cloud on the device, for clouds of millions of particles (the reference draws its clouds
 * serially, penning_trap.cc:51-161; am_set_particle_* reproduce those draws on the host).  Same
 * distributions, counter-based draws keyed on (seed, particle index), so NOT bit-identical to the
 * host generator: counts[s] particles of species s in enum order, positions uniform in the
 * ellipsoid of the given radii, velocities Maxwellian at temperatures[s] with each species' mean
 * removed and rescaled by sqrt(n / (n - 1)) as the reference does.  Fills the store (species
 * defaults for charge and mass, nothing annihilated, trap 0) and sets its particle count. */
int am_store_generate_cloud(am_store* store, const uint64_t counts[6], const double radii[3],
                            const double temperatures[6], uint64_t seed);

/* ---- context ------------------------------------------------------------------------------- */
int am_ctx_create(int device, am_ctx** out);
int am_ctx_destroy(am_ctx* ctx);
int am_ctx_synchronize(am_ctx* ctx);
/* Kernel timing: when enabled, every solver/push launch is bracketed by HIP events on the
 * context's stream; am_ctx_kernel_time returns the summed milliseconds and launch count of
 * the named kernel family ("allpairs_f32", "allpairs_f64", "push", "bh_walk", ...). */
int am_ctx_enable_timing(am_ctx* ctx, int enable);
int am_ctx_kernel_time(am_ctx* ctx, const char* name, double* total_ms, int64_t* launches);
int am_ctx_reset_timing(am_ctx* ctx);

/* ---- particle store -------------------------------------------------------------------------- */
int am_store_create(am_ctx* ctx, int64_t capacity, am_store** out);
int am_store_destroy(am_store* store);
/* Upload n particles (ChargedAntiparticle ctor, src/agents/charged_antiparticle.cc:5-15).
 * pos/vel: 3n doubles (x0,y0,z0,x1,...).  species: n bytes.  charge/mass: n doubles or NULL
 * (NULL = species defaults).  annihilated: n bytes or NULL.  trap: n uint32 or NULL (trap id
 * for ensembles, NULL = all 0).  Resets the Coulomb force to zero and the Lorentz
 * back-stagger flags (lorentz.h:34) to "not yet staggered". */
int am_store_upload(am_store* store, int64_t n, const double* pos, const double* vel, const uint8_t* species,
                    const double* charge, const double* mass, const uint8_t* annihilated, const uint32_t* trap);
/* Download any subset (NULL skips): pos/vel/force 3n doubles, annihilated/staggered n bytes. */
int am_store_download(am_store* store, double* pos, double* vel, double* coulomb_force, uint8_t* annihilated,
                      uint8_t* staggered);
/* Overwrite the Coulomb force buffer (SetCoulombForce for all agents); 3n doubles. */
int am_store_set_coulomb_force(am_store* store, const double* force);
int am_store_count(am_store* store, int64_t* n, int64_t* n_active);
/* The species of the n particles (GetSpecies for every agent): n bytes. */
int am_store_download_species(am_store* store, uint8_t* species);

/* ---- Coulomb solvers: write the store's Coulomb force for every alive particle ---------------- */
/* ComputeStandardCoulombForces (src/interaction/coulomb_interaction.h:10, .cc:23-80): pairs with
 * r^2 < r_c^2, cell list of cell size r_c.  cutoff_radius <= 0 sets all forces to zero.  Targets
 * with a non-finite coordinate get zero force (coulomb_interaction.cc:55-56).  precision: AM_FP64
 * is the reference arithmetic; AM_FP32 sums pairs in FP32 on hi/lo coordinates (FP64 per stencil
 * cell) with the same pair set: r^2 within 1e-5 relative of r_c^2 is decided in FP64. */
int am_compute_standard_coulomb_forces(am_store* store, double cutoff_radius, int precision);
/* Exact all-pairs force without cutoff: the theta = 0 limit of ComputeBarnesHutCoulombForces
 * (src/interaction/barnes-hut.h:22-24) with the AddPointCharge pair term (barnes-hut.cc:19-27). */
int am_compute_allpairs_coulomb_forces(am_store* store, int precision);
/* All-pairs FP32 forces that allocate nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The arithmetic is that of am_compute_allpairs_coulomb_forces(AM_FP32)
 * below 4,096 particles (the unsorted tile kernel, bit-identical there); the inputs and output are
 * caller-owned DEVICE buffers: pos 3n doubles (x0,y0,z0,x1,...) in m, charge n doubles in C,
 * annihilated n bytes or NULL, force 3n doubles out in N (zero for annihilated or neutral
 * particles).  Work is queued on `stream` (a hipStream_t; NULL = the null stream) of the current
 * device, and the call returns without synchronising.  Scratch is a 16 MiB file-scope device array
 * per GPU, held until the process ends; n above AM_NOALLOC_MAX_N returns AM_ERR_CAPACITY before any
 * launch.  Calls that overlap on one GPU share that array and race: serialise them. */
#define AM_NOALLOC_MAX_N 65536
int am_allpairs_forces_noalloc(int64_t n, const double* pos, const double* charge, const uint8_t* annihilated,
                               double* force, void* stream);
/* ComputeBarnesHutCoulombForces (src/interaction/barnes-hut.h:76, .cc:271-299): dual-monopole
 * octree, leaf capacity 8, MAC (2h)^2 < theta^2 |x - c|^2 against the geometric centre. */
int am_compute_barnes_hut_coulomb_forces(am_store* store, double theta, int precision);
/* The standalone BarnesHutTree (src/interaction/barnes-hut.h:28-74): Build over the store's alive
 * particles (leaf capacity 8), then ComputeForce(position, charge, self_index) and
 * ComputePotential(position, self_index) at n_points points given on the host: pos 3n doubles,
 * charge n doubles (NULL: zero charges, so zero force), self_index n int64 (NULL or -1: none; else
 * the store particle left out of the sum).  Outputs (host, either may be NULL): force 3n in N,
 * potential n in C/m (the sum of q_j / r_j, without the Coulomb constant).  AM_FP32 walks with the
 * FP32 pair math of am_compute_barnes_hut_coulomb_forces. */
int am_barnes_hut_evaluate(am_store* store, double theta, int precision, int64_t n_points, const double* pos,
                           const double* charge, const int64_t* self_index, double* force, double* potential);
/* ComputeBarnesHutCoulombForces plus the per-particle interaction count of BarnesHutTree::Accumulate
 * (barnes-hut.cc:183-221): leaf pairs + accepted monopoles, n int64 into counts (host). */
int am_barnes_hut_forces_counted(am_store* store, double theta, int precision, int64_t* counts);
/* BarnesHutTree::NumNodes (barnes-hut.h:46) and build statistics of the most recent Barnes-Hut build:
 * alive particles, nodes, deepest level, and depth-21 leaves holding more than 8 particles (the
 * reference, whose depth cap is 32, would split those further). */
int am_bh_tree_stats(am_store* store, int64_t* n_alive, int64_t* n_nodes, int32_t* max_depth,
                     int64_t* deep_overflow);
/* BarnesHutTree::order_ and nodes_ (barnes-hut.h:49-73) of the most recent build, in the reference's
 * DFS preorder: order[n_alive] particle indices; per node begin/end/is_leaf and the two monopole
 * charges (any pointer may be NULL). */
int am_bh_tree_download(am_store* store, uint32_t* order, int32_t* begin, int32_t* end, uint8_t* leaf, double* qp,
                        double* qn);
/* ComputeFastMultipoleCoulombForces (src/interaction/fast-multipole.h:73, .cc:684-713). */
int am_compute_fast_multipole_coulomb_forces(am_store* store, int order, double theta);
/* ComputeFastMultipoleCoulombForces with a choice of arithmetic: AM_FP64 is the reference arithmetic
 * (= am_compute_fast_multipole_coulomb_forces).  AM_FP32 evaluates P2P on hi/lo FP32 coordinates
 * with FP64 accumulation per work item and, for orders <= 5, the M2L contraction on FP32 matrix
 * cores (v_mfma_f32_16x16x4f32) with FP32 multipoles and FP32 derivative tensors; the tree, P2M,
 * M2M, L2L and L2P stay FP64. */
int am_compute_fast_multipole_coulomb_forces_ex(am_store* store, int order, double theta, int precision);
/* Statistics of the most recent FMM call on this store (the sizes FastMultipoleSolver::Compute builds,
 * fast-multipole.cc:397-526): sources, cells, M2L cell pairs, P2P cell pairs, deepest level, and the
 * number of distinct M2L displacement classes (any pointer may be NULL). */
int am_fmm_stats(am_store* store, int64_t* n_sources, int64_t* n_cells, int64_t* m2l_pairs,
                 int64_t* p2p_cell_pairs, int32_t* max_depth, int64_t* m2l_classes);
/* More statistics of the most recent FMM call (no reference counterpart): M2L pairs whose
 * displacement fits no class (they get tensors of their own), and the work items of the FP32 MFMA
 * M2L (0 when the FP64 contraction ran). */
int am_fmm_m2l_stats(am_store* store, int64_t* unclassed, int64_t* items);

/* ---- behaviours ---------------------------------------------------------------------------- */
/* LorentzForce::Run for every alive agent (src/forces/lorentz.cc:37-71): F = qE + F_coul,
 * first-call back-stagger v <- Boris(v, F, -dt/2), v <- Boris(v, F, dt), x += v dt, F_coul <- 0. */
int am_lorentz_force(am_store* store, const am_trap_field* field, double dt);
/* CyclotronCooling::Run for every alive agent (src/forces/cooling.cc:24-65).  The Langevin
 * noise (bath_temp_K > 0) draws counter-based normals keyed on (seed, step, particle index). */
int am_cyclotron_cooling(am_store* store, const am_trap_field* field, double tau_scale, double bath_temp_K,
                         double dt, uint64_t seed, uint64_t step);
/* LorentzForce then CyclotronCooling fused in one kernel (the per-agent behaviour order of
 * src/antimatter-kernel.cc:93-94); tau_scale <= 0 disables cooling (cooling.cc:35). */
int am_push(am_store* store, const am_trap_field* field, double dt, double tau_scale, double bath_temp_K,
            uint64_t seed, uint64_t step);
/* AnnihilationInteraction::Run for every agent (src/interaction/annihilation.cc:21-61) in the
 * canonical order: on current positions, scan i ascending; candidates j ascending (alive,
 * opposite matter per IsAnnihilationPair, r^2 < r_a^2); the first candidate whose draw
 * U(seed, step, i, j) < p marks both.  n_annihilated (optional) receives the number marked. */
int am_annihilation(am_store* store, double contact_radius, double probability, uint64_t seed, uint64_t step,
                    int64_t* n_annihilated);
/* IsAntiparticlePair (src/interaction/annihilation.h:9, .cc:12-19). */
int am_is_antiparticle_pair(int species_a, int species_b);

/* ---- energy ---------------------------------------------------------------------------------- */
/* ComputeTotalEnergy (src/analysis/energy_analysis.h:35, .cc:84-145): KE = sum m v^2 / 2,
 * PE_trap = sum q phi(x), PE_coul by settings->method (standard: shifted cutoff potential,
 * r <= r_c; BH / FMM potential; all-pairs), FP64 reductions.  history (optional) time-centres
 * the potentials with the previous call. */
int am_compute_total_energy(am_store* store, const am_trap_field* field, const am_energy_settings* settings,
                            am_energy_history* history, am_energy_breakdown* out);
/* ComputeTotalEnergy with a choice of pair arithmetic for the Coulomb potential: AM_FP64 is
 * am_compute_total_energy; AM_FP32 evaluates the all-pairs potential on the sorted FP32 tiles and the
 * Barnes-Hut potential with the FP32 walk (hi/lo coordinates, FP32 per tile or leaf, FP64 across them
 * and in every reduction); pairs under the 1e-15 m floor are skipped in both.  The standard and FMM
 * potentials are FP64 in both modes. */
int am_compute_total_energy_ex(am_store* store, const am_trap_field* field, const am_energy_settings* settings,
                               am_energy_history* history, int precision, am_energy_breakdown* out);

/* ---- batched ensemble of independent traps (no single reference counterpart: each trap is one
 * reference Simulate loop, src/antimatter-kernel.cc:153-197, with its own field and settings) ---- */
typedef struct am_ensemble am_ensemble;
typedef struct {
  int64_t begin, end;          /* particle range [begin, end) of the trap in the store */
  double b0, trap_voltage, d;  /* PenningTrapField */
  double dt;                   /* Lorentz step (PhysicsParams dt rule per trap) */
  double tau_scale, bath_temp; /* CyclotronCooling (tau_scale <= 0: off) */
  int32_t annihilation;        /* AnnihilationInteraction attached */
  double contact_radius, probability;
} am_trap_desc;
/* Creates an ensemble over ascending, disjoint trap ranges of a store (writes the store's trap ids).
 * contact_capacity: candidate-pair buffer of the contact step (<= 0: 16 per particle).  Per step:
 * all-pairs forces inside each trap (AM_FP32 arithmetic), LorentzForce + CyclotronCooling,
 * AnnihilationInteraction in canonical order.  Draws of trap t are keyed by
 * seed + t * 0x9E3779B97F4A7C15 and trap-local particle indices. */
int am_ensemble_create(am_store* store, int32_t n_traps, const am_trap_desc* traps, int64_t contact_capacity,
                       am_ensemble** out);
int am_ensemble_destroy(am_ensemble* ens);
/* Source-chunk length of the force work items (>= 128). */
int am_ensemble_set_chunk(am_ensemble* ens, int32_t chunk);
/* Autotuner: times each candidate chunk length (NULL: 1024..32768) on the device, keeps the fastest. */
int am_ensemble_autotune(am_ensemble* ens, const int32_t* candidates, int32_t n_candidates, int32_t* chosen,
                         double* times_ms);
/* Advances every trap `steps` steps; use_graph replays one hipGraph-captured step. */
int am_ensemble_run(am_ensemble* ens, int32_t steps, uint64_t seed, int32_t use_graph);
int am_ensemble_stats(am_ensemble* ens, int64_t* steps_done, int64_t* annihilated, int32_t* chunk, int64_t* n_items,
                      double* pair_interactions_per_step);

/* ---- learned far-field correction of Barnes-Hut (no reference counterpart: it corrects the
 * dual-monopole approximation of the walk in src/interaction/barnes-hut.cc:183-221) -------------
 * For each accepted node and charge sign s the walk's monopole error is predicted by a small MLP
 * (16 -> 64 -> 64 -> 3, SiLU; BF16 weights on MFMA, FP32 accumulation) from 13 features of the
 * node's sign-s charges: r/R (3), h/R, (c_s - c_geo)/h (3), S_s/(Q_s h^2) (6), where r = target -
 * charge centre, R = |r|, h = half width, S_s the second moments about c_s.  Output y is the
 * residual in units of sigma_s Q_s h^2 / R^4 (Coulomb constant and target charge factored out). */
typedef struct am_mlp am_mlp;
/* New network on the context's device; weights drawn from seed (He normal). */
int am_mlp_create(am_ctx* ctx, uint64_t seed, am_mlp** out);
int am_mlp_destroy(am_mlp* net);
/* Number of FP32 parameters (W1 16x64, b1, W2 64x64, b2, W3 64x3, b3; row-major [in][out]). */
int am_mlp_num_params(void);
int am_mlp_set_params(am_mlp* net, const float* params);
int am_mlp_get_params(am_mlp* net, float* params);
/* Prediction mode: 0 (the default) predicts y with the network alone; quadrupole != 0 predicts
 * y = y_q(features) + network, where y_q = -(Qt u - 5/2 (u.Qt.u) u), Qt = 3 S - tr(S) I, is the
 * analytic quadrupole term and the network is trained on y - y_q. */
int am_mlp_set_baseline(am_mlp* net, int quadrupole);
/* The analytic quadrupole term y_q of raw feature rows X[rows][16] -> Y[rows][3] (host). */
int am_mlp_quadrupole_baseline(const float* X, int64_t rows, float* Y);
/* Input standardisation x' = (x - mean) / std (16 each) and output scale y = out_scale * y' (3). */
int am_mlp_set_normalization(am_mlp* net, const float* in_mean, const float* in_std, const float* out_scale);
/* Mean-squared-error gradient of a batch of normalised inputs X[rows][16] against normalised
 * targets Y[rows][3] (forward with BF16-rounded weights and activations): grad[num_params], loss. */
int am_mlp_gradients(am_mlp* net, const float* X, const float* Y, int32_t rows, float* grad, float* loss);
/* Adam training (lr, beta 0.9/0.999) on raw features X[rows][16] and labels Y[rows][3] with the
 * network's normalisation (targets (Y - y_q) / out_scale in the quadrupole mode): `epochs` passes of
 * `batch`-row minibatches in a seeded shuffled order; losses (optional) receives the mean normalised
 * loss per epoch. */
int am_mlp_train(am_mlp* net, const float* X, const float* Y, int64_t rows, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float* losses);
/* MFMA inference on raw features X[rows][16] -> predictions Y[rows][3] (host arrays). */
int am_mlp_infer(am_mlp* net, const float* X, int32_t rows, float* Y);
/* Training rows from the store's cloud: Barnes-Hut walk at theta over n_targets particles (every
 * stride-th in the tree order from offset), up to rec_cap accepted nodes each, one row per node and
 * charge sign present: raw features X[16], the exact residual in units of the label scale Y[3],
 * and, each optional (NULL skips), the target particle index, the node (DFS preorder position, as
 * in am_bh_tree_download) and the label scale sigma_s Q_s h^2 / R^4 (e / (2^-20 m)^2).  Outputs
 * hold 2*n_targets*rec_cap rows; *rows_out receives the number written. */
int am_bh_learned_training_data(am_store* store, double theta, int32_t n_targets, int32_t stride, int32_t offset,
                                int32_t rec_cap, float* X, float* Y, int32_t* target, int32_t* node, float* scale,
                                int64_t* rows_out);
/* ComputeBarnesHutCoulombForces (AM_FP32 walk) plus the learned correction of every accepted node.
 * A row whose |y| exceeds ybound or whose h/R exceeds amax takes the exact residual instead.
 * rec_cap: nodes recorded per particle (later ones keep the plain monopole).  stats (optional, 3):
 * rows corrected by the model, rows that fell back, particles whose node list overflowed. */
int am_compute_barnes_hut_learned_forces(am_store* store, double theta, am_mlp* net, int32_t rec_cap, float ybound,
                                         float amax, int64_t* stats);

/* ---- diagnostics (no reference counterpart: checks of the device building blocks) ------------ */
/* Exclusive prefix sum of n uint32 on the device; out receives n + 1 values (out[n] = total). */
int am_debug_exclusive_scan(am_ctx* ctx, const uint32_t* in, int64_t n, uint32_t* out);
/* Stable LSD radix sort of (key, value) pairs on the device, in place on the host arrays. */
int am_debug_radix_sort(am_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits);
/* Stable radix sort of keys alone on bits [bit0, bit0 + key_bits) on the device, in place on the host
 * array; n < 2^30. */
int am_debug_radix_sort_keys(am_ctx* ctx, uint64_t* keys, int64_t n, int bit0, int key_bits);
/* FMM M2L options of the context (defaults 1 and 2^20): classes = 0 gives every M2L pair a tensor of
 * its own (no displacement classes); max_unclassed is the most unclassed pairs that the FP32 MFMA
 * M2L takes, past which the FP64 contraction runs (negative: always the FP64 contraction). */
int am_debug_fmm_m2l_options(am_ctx* ctx, int classes, int64_t max_unclassed);

#ifdef __cplusplus
}
#endif

#endif /* ANTIMATTER_AMD_AM_API_H_ */
