/* pgm_hip: pGM electrostatics on one AMD MI355X (gfx950), C ABI.
 *
 * The engine of the reference's core/cuda/pgm_cuda.h (pgmcuda::Engine, core/cuda/pgm_cuda.h:47-127),
 * re-designed for CDNA4 and exposed as plain C: an opaque handle, plain pointers and sizes, int status
 * (0 = success, nonzero = error, message in pgm_last_error), no exceptions across the boundary.
 *
 * What it computes (as core/pgm_core.cpp does on the CPU): Gaussian charges, covalent permanent dipoles and
 * Gaussian induced dipoles with smooth PME; the induced dipoles by Jacobi-preconditioned CG with the mu4
 * predictor; energy terms, forces and the strain derivative S_ab = dE/d eps_ab.
 *
 * Precision: PGM_DOUBLE runs every kernel in FP64.  PGM_MIXED runs the pair and PME arithmetic in FP32 with
 * FP64 accumulation of fields, forces, energies and the strain, and keeps the CG vectors, dot products and
 * convergence norms in FP64 (FP32 matrix-vector products inside an FP64 Krylov recurrence).
 *
 * Units: nm, e, e nm, nm^3, kJ/mol.  Box: rows a, b, c, reduced triclinic (a = (ax,0,0), b = (bx,by,0)),
 * row-major double[9].  Forces and the strain are ADDED to the caller's arrays.  Threading: one thread at a
 * time per engine (any thread), as core/cuda/pgm_cuda.h:18.
 */
#ifndef PGM_HIP_H_
#define PGM_HIP_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgm_engine pgm_engine;

enum { PGM_DOUBLE = 0, PGM_MIXED = 1 };

/* pgmcuda::Settings (core/cuda/pgm_cuda.h:25-39) plus precision and polarizable
 * (pgmcore::Settings::polarizable, core/pgm_core.h:27). */
typedef struct {
    double cutoff;    /* direct-space cutoff [nm] */
    double beta;      /* Ewald coefficient [1/nm] */
    int grid[3];      /* PME grid */
    int order;        /* B-spline order, 4..8 */
    double tol;       /* max|alpha r| / mean|alpha b| */
    int max_iter;
    double peek;      /* final mu += peek alpha r */
    int predictor;    /* 0 none, 1 mu3, 2 mu4 */
    double ke;        /* Coulomb constant [kJ mol^-1 nm e^-2] */
    int device;       /* HIP device; < 0: the device current on the calling thread, never switched */
    double skin;      /* pair list: cutoff + skin, rebuilt when an atom moved > skin/2 [nm] */
    int precision;    /* PGM_DOUBLE or PGM_MIXED */
    int polarizable;  /* 1: induced dipoles; 0: permanent multipoles only */
} pgm_settings;

/* pgmcuda::Terms (core/cuda/pgm_cuda.h:41-45), kJ/mol. */
typedef struct {
    double direct, reciprocal, self, background, polarization, gvdw;
} pgm_terms;

/* Defaults of pgmcuda::Settings (core/cuda/pgm_cuda.h:25-39) with pgmcore's Coulomb constant 138.935456
 * (core/pgm_core.h:33) and PGM_MIXED. */
void pgm_default_settings(pgm_settings* s);

/* Engine(const Settings&) (core/cuda/pgm_cuda.h:49).  NULL on failure (message: pgm_last_error(NULL)). */
pgm_engine* pgm_create(const pgm_settings* s);
/* ~Engine (core/cuda/pgm_cuda.h:50). */
void pgm_destroy(pgm_engine* e);
/* The last error of e (NULL: of the last failed pgm_create on this thread). */
const char* pgm_last_error(const pgm_engine* e);

/* setAtoms (core/cuda/pgm_cuda.h:55-56): n atoms in the caller's order; global_index (NULL: i) keys the
 * predictor history (n_global slots). */
int pgm_set_atoms(pgm_engine* e, int n, const double* q, const double* radius, const double* alpha,
                  const int* global_index, int n_global);
/* Engine::setGvdw (core/cuda/pgm_cuda.h:57-65): GVDW, the Gaussian-density van der Waals of core/gvdw_kernel.h,
 * by the combination rule (A_ij = a_i a_j, C6_ij = c_i c_j, B_ij = (b_i + b_j)/2 from per-atom sqrt_a, sqrt_c6,
 * b; zeros: none) for all pairs within rvdw except the skipped pairs (skip_i[k], skip_j[k]), each given once
 * (exclusions, and pairs with their own parameters: those and the dispersion tail are the caller's).
 * repulsion: 0 Gaussian, 1 Slater.  Its energy is pgm_terms.gvdw [kJ/mol]; its forces and strain are added to
 * the electrostatic ones.  After every pgm_set_atoms and pgm_set_batch; rvdw = 0 turns it off.  Unlike the reference, rvdw may
 * not exceed the electrostatic cutoff (the pair list radius), and pgm_backward refuses an engine with GVDW on. */
int pgm_set_gvdw(pgm_engine* e, int repulsion, double rvdw, const double* sqrt_a, const double* sqrt_c6,
                 const double* b, int nskip, const int* skip_i, const int* skip_j);
/* Engine::setCandidates (core/cuda/pgm_cuda.h:68-73): the caller's pair candidates instead of the engine's own
 * list, until the next pgm_set_atoms.  Segment s (nseg of them) holds the pairs (seg_i[s], cand_j[k]),
 * k in [seg_begin[s], seg_begin[s+1]) (seg_begin: nseg + 1 entries), with x = r_i + t.box - r_j for the integer
 * image t = seg_shift[3s..3s+2] (|t_k| <= 7); every pair at most once; pairs beyond the cutoff are skipped at
 * every evaluation.  Positions are used as given (not wrapped into the box), the list is never rebuilt, and the
 * engine keeps the caller's atom order.  One system of at most 2^20 atoms; not with pgm_set_batch or GVDW. */
int pgm_set_candidates(pgm_engine* e, int nseg, const int* seg_i, const int* seg_shift, const int* seg_begin,
                       const int* cand_j);
/* The CG preconditioner, the reference CUDA engine's PGM_CUDA_PRECOND and PGM_CUDA_JACOBI_SELF options
 * (core/cuda/pgm_cuda.cu:1753-1760, precondK :1526-1588): radius > 0 [nm, at most the cutoff] applies the
 * short-range M = D + D T_s D, D = diag(alpha), T_s the direct-space dipole tensor of the pairs within radius
 * (from the stored pair kernels; Jacobi everywhere when an atom has more than 48 of them); radius 0 is Jacobi,
 * with jacobi_self != 0 alpha/(1 - alpha c_self).  The tolerance stays on max |alpha r|; the solution is the same,
 * the iteration count changes.  Unlike the reference, which uses it only when the predicted guess rides on the
 * pair setup, it applies to every solve, the adjoint's included.  Kept across pgm_set_atoms; default Jacobi. */
int pgm_set_preconditioner(pgm_engine* e, double radius, int jacobi_self);
/* setCovalent (core/cuda/pgm_cuda.h:67): p_atom += strength unit(r_partner - r_atom). */
int pgm_set_covalent(pgm_engine* e, int ncov, const int* atom, const int* partner, const double* strength);

/* compute (core/cuda/pgm_cuda.h:79): host positions (3n), forces added to f (3n, may be NULL), strain
 * added to strain (9, may be NULL), terms out (may be NULL).  new_step: this evaluation starts a new MD step
 * (its solution enters the predictor history). */
int pgm_compute(pgm_engine* e, const double* pos, const double* box, int new_step, double* f, double* strain,
                pgm_terms* terms);
/* computeF (core/cuda/pgm_cuda.h:82): float positions; forces kept for pgm_forces. */
int pgm_compute_f32(pgm_engine* e, const float* pos, const double* box, int new_step, double* strain,
                    pgm_terms* terms);
/* forces() (core/cuda/pgm_cuda.h:86): the forces of the last pgm_compute_f32 (3n, kJ/mol/nm). */
int pgm_forces(pgm_engine* e, double* out);

/* computeDevice (core/cuda/pgm_cuda.h:93-94): positions from the device float4 array posq4 (w ignored,
 * caller's order), forces added atomically to the device fixed-point buffer force_fixed (x of atom i at i,
 * y at i + stride, z at i + 2 stride, scale fixed_scale); stream: the caller's hipStream_t (0: null stream),
 * whose work so far comes first. */
int pgm_compute_device(pgm_engine* e, const void* posq4, const double* box, int new_step, long long* force_fixed,
                       int stride, double fixed_scale, void* stream, pgm_terms* terms);
/* launchDevice / finishDevice (core/cuda/pgm_cuda.h:95-104): pgm_compute_device in two parts, the caller doing
 * other work on the GPU in between.  pgm_launch_device enqueues the whole evaluation behind the work so far on
 * stream and returns without waiting for the GPU: the pair list is checked on the device, the host predicting
 * from the last step's displacement and largest move whether it will hold (rebuilt before the launch if not);
 * the CG iterations enqueued are the most the last 16 solves needed (one more if their counts varied), each
 * skipped on the device once the solve has converged.  Exceptions that wait: a list build, and the first launch of an engine or a repeat after a
 * short solve, whose CG checks convergence on the host between chunks of iterations.  The forces are added to
 * force_fixed in every case.  pgm_finish_device waits for the launch and sets valid = 0 when its list turned
 * out to be stale for these positions (the list is rebuilt now) or its solve needed more iterations than were
 * enqueued: the forces added are then not valid, and the caller repeats the launch with the same positions and
 * new_step (the same guess and history entry: an invalid launch does not advance the history). */
int pgm_launch_device(pgm_engine* e, const void* posq4, const double* box, int new_step, long long* force_fixed,
                      int stride, double fixed_scale, void* stream);
int pgm_finish_device(pgm_engine* e, pgm_terms* terms, int* valid);
/* waitDevice (core/cuda/pgm_cuda.h:105-111): instead of pgm_finish_device, the caller's stream waits on the
 * GPU for the forces of the launch (no host wait).  Its completion happens at the next call of any kind: a
 * stale list or a short solve found then is counted by pgm_late_list_checks (the forces were used). */
int pgm_wait_device(pgm_engine* e, void* stream);
/* No counterpart in the reference: 1 while the GPU work of a launch is still running, 0 when it has completed
 * or no launch is in flight (a poll without waiting, for callers that schedule other host work). */
int pgm_launch_pending(const pgm_engine* e);

/* No counterpart in the reference: compute on FP64 device arrays, for callers that keep the system in HBM
 * (positions 3n doubles in the caller's order; forces added to f_dev, 3n doubles, may be NULL). */
int pgm_compute_device_f64(pgm_engine* e, const double* pos_dev, const double* box, int new_step, double* f_dev,
                           double* strain, void* stream, pgm_terms* terms);

/* No counterpart in the reference; for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): pgm_compute_device_f64 (new_step 1) that
 * makes no hipMalloc, hipFree or other allocation.  Every buffer of the evaluation is carved from the library's
 * static device scratch of 128 MiB (one per GPU, held until the process ends), so the engine holds at most 4096
 * atoms (and a pair list that fits; an error says when it does not).  Nothing outlives the call: the pair list is
 * built, the solve starts from alpha E (the predictor history is neither read nor written).  The calls on one
 * GPU share its scratch, which keeps the last call's contents: calls that overlap on one GPU race.  One system
 * (not a batch), the engine's own pair list.  The setup calls (pgm_create, pgm_set_*) allocate as usual. */
int pgm_compute_device_f64_noalloc(pgm_engine* e, const double* pos_dev, const double* box, double* f_dev,
                                   double* strain, void* stream, pgm_terms* terms);

/* pgmcore::Engine::setGeometry + field (core/pgm_core.h:76-79): the field [e/nm^2] at every atom (field, 3n,
 * caller's order, assigned) of the charges q (n) and dipoles d (3n, e nm) given here, direct + reciprocal + self,
 * at host positions pos (3n) and box; the engine's own charges, covalent and induced dipoles play no part.  The
 * pair list is checked (rebuilt if stale) as for an evaluation; the results of the last evaluation are gone. */
int pgm_field(pgm_engine* e, const double* pos, const double* box, const double* q, const double* d, double* field);

/* iterations, residual, inducedDipoles, permanentDipoles, resetHistory, listBuilds, numCandidates,
 * lateListChecks (core/cuda/pgm_cuda.h:112-122). */
int pgm_iterations(const pgm_engine* e);
double pgm_residual(const pgm_engine* e);
int pgm_induced_dipoles(pgm_engine* e, double* mu);
int pgm_permanent_dipoles(pgm_engine* e, double* p);
int pgm_reset_history(pgm_engine* e);
long pgm_list_builds(const pgm_engine* e);
long long pgm_num_candidates(const pgm_engine* e);
long pgm_late_list_checks(const pgm_engine* e);
/* No counterpart in the reference: a test hook.  out[5]: what the next pgm_launch_device predicts the list's
 * staleness from, the largest displacement since the build at the last completed check [nm], the largest move
 * of the last step [nm] and the skin [nm]; the launches repeated so far; the synchronous evaluations redone
 * because their solve needed more iterations than were enqueued ahead. */
int pgm_list_check_state(const pgm_engine* e, double* out);
/* No counterpart in the reference: pairs within the cutoff at the last evaluation, each counted once. */
long long pgm_num_pairs(pgm_engine* e);

/* No counterpart in the reference: load nset prior solutions (each 3*n_global doubles, global order, most
 * recent first) as the predictor history, e.g. the converged dipoles of earlier steps of a trajectory. */
int pgm_set_history(pgm_engine* e, int nset, const double* mu_sets);
/* No counterpart in the reference: freeze (1) or unfreeze (0) the history, so that repeated evaluations
 * of one workload start from the same predicted guess (benchmarks). */
int pgm_freeze_history(pgm_engine* e, int frozen);

/* No counterpart in the reference (it has no parameter derivatives; its parameters come from the prmtop
 * sections read by openmm/python/pgm_prmtop.py:54-115): gradients of a loss L(E, F) with respect to the
 * per-atom charges, radii and polarizabilities and the per-entry covalent strengths, for force-matching fits.
 * Evaluates at pos (as a same-step re-evaluation: the history is not changed), then solves the adjoint
 * induction system with the same symmetric matrix and CG as the forward solve (tolerance tol) and recomputes
 * the PME grids.  dl_dforces: dL/dF (3n doubles, caller's order; NULL for an energy-only loss); dl_denergy:
 * dL/dE.  Outputs are assigned (not added), caller's order: grad_q, grad_radius, grad_alpha (n doubles each),
 * grad_cov (ncov doubles, the order of pgm_set_covalent); any may be NULL.  Units: kJ/mol per e, nm, nm^3,
 * e nm times the units of dL/dE (per kJ/mol) and dL/dF (per kJ/mol/nm). */
int pgm_backward(pgm_engine* e, const double* pos, const double* box, const double* dl_dforces, double dl_denergy,
                 double* grad_q, double* grad_radius, double* grad_alpha, double* grad_cov);
/* The same with device positions, cotangent and outputs, ordered after the work on stream (a hipStream_t;
 * NULL: the null stream). */
int pgm_backward_device(pgm_engine* e, const double* pos_dev, const double* box, const double* dl_dforces_dev,
                        double dl_denergy, double* grad_q_dev, double* grad_radius_dev, double* grad_alpha_dev,
                        double* grad_cov_dev, void* stream);
/* CG iterations and final residual of the last adjoint solve. */
int pgm_adjoint_iterations(const pgm_engine* e);
double pgm_adjoint_residual(const pgm_engine* e);

/* No counterpart in the reference (it evaluates one system per engine): a batch of independent periodic
 * systems packed into one engine, for parameter fitting over many snapshots of uneven size.  Atoms
 * [atom_offsets[s], atom_offsets[s+1]) of those set by pgm_set_atoms form system s (atom_offsets[0] = 0,
 * atom_offsets[nsys] = n), with PME grid grids[3s..3s+2] (NULL: the settings' grid for all).  Covalent
 * dipoles may not join two systems.  Every pass of an evaluation is one launch over all systems (the FFTs one
 * per grid shape over the stacked grids of that shape), and one CG solves the block-diagonal induction system.
 * The single-system entry points refuse a batch of more than one system; pgm_set_atoms with another atom count
 * resets the engine to one system. */
int pgm_set_batch(pgm_engine* e, int nsys, const int* atom_offsets, const int* grids);
int pgm_num_systems(const pgm_engine* e);
/* One evaluation of every system of the batch: boxes 9 doubles per system (reduced triclinic rows); forces
 * added to f (3n doubles, caller's order, may be NULL); terms one per system (may be NULL). */
int pgm_compute_batch(pgm_engine* e, const double* pos, const double* boxes, int new_step, double* f,
                      pgm_terms* terms);
int pgm_compute_batch_device(pgm_engine* e, const double* pos_dev, const double* boxes, int new_step, double* f_dev,
                             void* stream, pgm_terms* terms);
/* pgm_backward over a batch: dl_denergy holds dL/dE per system (NULL: zero). */
int pgm_backward_batch(pgm_engine* e, const double* pos, const double* boxes, const double* dl_dforces,
                       const double* dl_denergy, double* grad_q, double* grad_radius, double* grad_alpha,
                       double* grad_cov);
int pgm_backward_batch_device(pgm_engine* e, const double* pos_dev, const double* boxes,
                              const double* dl_dforces_dev, const double* dl_denergy, double* grad_q_dev,
                              double* grad_radius_dev, double* grad_alpha_dev, double* grad_cov_dev, void* stream);

/* No counterpart in the reference: the autotuner of the pair passes.  After an evaluation, times the
 * compaction + setup, (iterations + 1) matrix-vector and final pair passes on the engine's current lists for
 * two variants (compaction fused into the setup pass, or as a pass of its own) and 16, 32 and 64 lanes per
 * atom (reps repetitions each), keeps the fastest for later evaluations and returns it in *lanes, *fused
 * (each may be NULL) and the six timings in ms[6] (ms per evaluation, fused 16/32/64 then separate
 * 16/32/64, each the faster of the two field-pass forms in FP32; may be NULL).  It then times the PME beside
 * or after the pair passes (pgm_set_pme_overlap) and keeps the faster too.  pgm_set_lanes sets the pair
 * passes directly (lanes 16, 32 or 64; 16 and separate passes by default). */
int pgm_autotune(pgm_engine* e, int reps, int* lanes, int* fused, double* ms);
int pgm_set_lanes(pgm_engine* e, int lanes, int fused);
/* No counterpart in the reference: the lanes per atom of the pair setup pass.  pgm_autotune, once it has chosen
 * the pair passes, times the setup pass alone at 16, 32 and 64 lanes and keeps the fastest for it (the setup
 * tests every candidate, the other passes stream the compacted pairs); pgm_set_lanes sets it with the others,
 * pgm_set_setup_lanes alone (16, 32 or 64, after pgm_set_lanes). */
int pgm_setup_lanes(const pgm_engine* e);
int pgm_set_setup_lanes(pgm_engine* e, int lanes);
/* No counterpart in the reference: the FP32 (PGM_MIXED) field pass of the CG products with its pair kernels
 * evaluated on the fly (1) or read from the kernels the setup pass stored (0); pgm_autotune times both and keeps
 * the faster.  pgm_field_otf: the current choice. */
int pgm_set_field_otf(pgm_engine* e, int on);
int pgm_field_otf(const pgm_engine* e);
/* No counterpart in the reference (whose PME solves run in line): the PME solves of an evaluation on a second
 * stream, with a hardware queue of its own, beside the pair pass they are independent of (1, the default), or
 * after it on the engine stream (0).  Results are the same bit for bit either way.  pgm_autotune times one CG
 * product (field pass + PME solve) both ways and keeps the faster; pgm_pme_overlap_times gives those two
 * timings, ms[0] after, ms[1] beside (ms per product; zeros before an autotune). */
int pgm_set_pme_overlap(pgm_engine* e, int on);
int pgm_pme_overlap(const pgm_engine* e);
int pgm_pme_overlap_times(pgm_engine* e, double* ms);

/* No counterpart in the reference (whose engine prints host wall clocks under PGM_CUDA_TIMING,
 * core/cuda/pgm_cuda.cu:1814-1817): HIP-event timing of the kernel groups of every evaluation, for
 * profiling.  Phases, in order: list check/rebuild, per-evaluation preparation, pair compaction, pair
 * setup, pair field (matrix-vector products), PME spread, PME transforms + convolution, PME gather, CG
 * vector kernels, final pair pass, final reductions and force output.  Iterations enqueued ahead that found the solve
 * finished are not counted.  pgm_set_timing resets the sums. */
enum { PGM_TIMING_PHASES = 11 };
int pgm_set_timing(pgm_engine* e, int on);
int pgm_get_timing(pgm_engine* e, double* ms, long* counts, int nphases);
/* As pgm_set_timing(e, mask != 0), bracketing only the phases of the mask (bit k: phase k): each bracket
 * records two events on the engine stream, which leave the GPU idle for a few microseconds, so a timed
 * production run brackets the one kernel group it measures. */
int pgm_set_timing_phases(pgm_engine* e, unsigned mask);

/* No counterpart in the reference: test hooks that evaluate the device's own math on host arrays (they
 * allocate, copy and synchronize on the current device; for tests, not for production loops).
 * pgm_eval_erfc_f32: y[i] = erfc(x[i]) as the FP32 pair kernels compute it (the Chebyshev fit of Press et al.,
 * Numerical Recipes 2nd ed. 6.2, with v_rcp_f32 and exp).  pgm_eval_pair_kernels: out[i][0..3] = G0..G3 of a
 * pair at squared distance r2[i] with rg_i + rg_j = rg_sum[i] (a = rg_sum^-1/2) and Ewald beta, as the pair
 * passes of the given precision compute them; out[i][4..5] = G0, G3 rebuilt from G1, G2 as the FP32 final and
 * tangent passes do (PGM_DOUBLE: G0, G3 again).  Errors: pgm_last_error(NULL). */
int pgm_eval_erfc_f32(int n, const float* x, float* y);
int pgm_eval_pair_kernels(int precision, int n, const double* rg_sum, const double* r2, double beta, double* out);

#ifdef __cplusplus
}
#endif

#endif
