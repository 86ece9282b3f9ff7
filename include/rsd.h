/*
 * rsd.h -- C ABI of the MI355X regularized Stokesian dynamics hot path.
 *
 * Drop-in boundary for the periodic FTS step of the Fortran reference
 * (sherlockyzd/Regularized-Stokesian-Dynamics, src/).  Every entry point
 * replaces one reference routine and is cited as file:line.  The Fortran
 * side binds these through iso_c_binding (fortran/rsd_gpu.f90); any C,
 * C++ or ctypes caller can use them directly.
 *
 * Conventions (as in the reference):
 *   - all arrays are caller-owned FP64 host memory, column-major (Fortran
 *     order); positions are CONF(3,N), i.e. x,y,z of particle 1, then 2, ...
 *   - grand vectors are field-major, particle-minor: translations rows
 *     3(a-1)+i, rotations 3N+3(a-1)+i, stresslet/strain 6N+5(a-1)+j
 *     (per_GRPY_YZD.f90:165-184); the 5-vector basis is tensors.f90:43-67.
 *   - mu_f is the (nondimensional) fluid viscosity; the reference forces
 *     it to 1/pi (mymodules.f90:816,846).
 *
 * Return codes: 0 ok; > 0 the Cholesky INFO of the first non-positive-
 * definite leading minor (non-fatal, as MATREV prints and continues,
 * matrices.f90:19-24); < 0 an RSD_ERR_* code, with rsd_last_error() text.
 * Nothing is printed.  One host thread per context; calls are blocking.
 */
#ifndef RSD_H_
#define RSD_H_

#ifdef __cplusplus
extern "C" {
#endif

#define RSD_OK 0
#define RSD_ERR_ARG (-1)
#define RSD_ERR_HIP (-2)
#define RSD_ERR_NOMEM (-3)
#define RSD_ERR_TABLES (-4)
#define RSD_ERR_CAPACITY (-5)
#define RSD_ERR_NODEVICE (-6)
#define RSD_ERR_UNSUPPORTED (-7)
#define RSD_ERR_NOT_SPD (-8)  /* a factorisation the reference stops on failed (CHOLESKY, matrices.f90:140-147) */

/* LATTICE_BASE module state (mymodules.f90:33-41). */
typedef struct rsd_lattice {
  double lr[9];  /* LR: lattice vectors as columns (column-major 3x3) */
  double li[9];  /* LI = LR^-1 */
  double lv;     /* |det LR| */
  double sig, lammda, logerr, phi; /* written by CALC_LATTICE_INVERSE */
  int nr, ni;                      /* real / reciprocal image half-widths */
} rsd_lattice;

/* Control settings the periodic FTS step reads (control_file.dat order,
 * initial_conf.f90:116-179; globals of mymodules.f90). */
typedef struct rsd_step_config {
  double lb[3];          /* LATTICE_BASE::LB box edges */
  double gamma;          /* SYS_property::GAMMA shear rate */
  double ews_err;        /* LINERR, the Ewald truncation error (EWS_ERR) */
  double mu_f;           /* fluid viscosity (1/pi after nondimensionalisation) */
  int correction_method; /* overlap correction: 0 none, 1 Yamakawa, 2 regularised */
  int nb;                /* SIZE::Nb boundary particles at the end of the list */
  double gravity;        /* SYS_property::gravity: body force -gravity e_z */
  double dt_dem0;        /* control::dt_dem0 collision sub-step scale */
  double radii_test;     /* CONFIG::radii_test (mean-velocity force scale) */
  double u_bg[3];        /* CONFIG::u_bg */
  double omega_bg[3];    /* CONFIG::omega_bg (component 2 is set to gamma/2 each step) */
  double ei_bg[6];       /* CONFIG::EI_bg (component 5 is set to gamma/2 each step) */
} rsd_step_config;

/* Fills a config with the defaults of the periodic FTS benchmark settings. */
void rsd_step_config_default(rsd_step_config* cfg, double box_edge);

const char* rsd_last_error(void);
int rsd_version(void);

/* Selects the HIP device used by the calling thread's subsequent calls. */
int rsd_device_init(int device);

/* Loads lubdat/r2babc.dat, r2bgh.dat, r2bm.dat from a directory
 * (replaces the first-call reads of hydro_lub.f90:455-466, 529-596). */
int rsd_lub_tables_load(const char* lubdat_dir);

/* SET_LATTICE_SKEW, csd_post.f90:105-136 (steady shear). */
int rsd_set_lattice_skew(double t, const double lb[3], double gamma, rsd_lattice* lat);
/* CALC_LATTICE_INVERSE, csd_post.f90:158-183: fills sig, lammda, logerr, nr, ni, phi. */
int rsd_calc_lattice_inverse(rsd_lattice* lat, double ews_err, int n, const double* radii);
/* PER_SKEW_CORRECTION, csd_post.f90:24-63 (host). */
int rsd_per_skew_correction(const rsd_lattice* lat, double r[3], double* dmin);

/* PER_GRPERY_INV_FRI, per_GRPY_YZD.f90:13-200.  lat must hold LR/LI/LV
 * (SET_LATTICE_SKEW); the Ewald constants are computed here and written back
 * into lat, as the reference writes SIG/Lammda/LOGERR/NR/NI/phi.
 * app(6N,6N), apq(6N,5N), aqq(5N,5N) out.  correction_method selects the
 * overlap correction of pairs with r <= a_i + a_j (per_GRPY_YZD.f90:98-116):
 * 0 none, 1 Yamakawa (overlap_Yamakawa_CORRECTION_IJ, GRPY_YZD.f90:789-811),
 * 2 regularised (overlap_Regular_CORRECTION_IJ, GRPY_YZD.f90:813-836).
 * Returns the number of overlap-corrected pairs in *n_overlap (may be NULL). */
int rsd_per_grpery_inv_fri(int n, const double* pos, const double* radii, rsd_lattice* lat, double ews_err,
                           int correction_method, double mu_f, double* app, double* apq, double* aqq,
                           int* n_overlap);

/* Allocation-free form of rsd_per_grpery_inv_fri.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Device pointers in, device matrix
 * out, stream-ordered on `stream` (a hipStream_t; NULL: the null stream), no
 * synchronisation and no allocation of any kind.
 *   d_pos(3,N), d_radii(N): device arrays.
 *   lat: the lattice with the Ewald constants already in it (rsd_set_lattice_skew
 *        then rsd_calc_lattice_inverse, or explicit nr/ni/lammda/phi); NI <= 7.
 *   d_grmobmx(ld, 11N), ld >= 11N: the whole grand mobility
 *        [APP APQ; APQ^T AQQ] of per_GRPY_YZD.f90:160-198 (already divided by
 *        pi mu_f), every entry written.
 *   d_n_overlap: device int receiving the count of overlap-corrected pairs,
 *        or NULL.
 * Its k-space weights live in a library-owned device array of 1.65 MiB per
 * device, resolved once per device (rsd_device_init, or the first call):
 * calls that overlap in time on one device race on it. */
int rsd_per_grpery_inv_fri_noalloc(int n, const double* d_pos, const double* d_radii, const rsd_lattice* lat,
                                   int correction_method, double mu_f, double* d_grmobmx, int ld, int* d_n_overlap,
                                   void* stream);

/* PER_GRPERY_INV_FRI with an explicit Ewald image set in place of the one the
 * reference derives from EWS_ERR (no counterpart in the reference): the real
 * space cube of half-width nr, the reciprocal cube of half-width ni, and with
 * kmax2 > 0 only the reciprocal vectors n with |n|^2 <= kmax2 (the benchmark
 * configurations' "4 real images, |k| <= 6 2pi/L" is nr = 4, ni = 6,
 * kmax2 = 36).  lambda and phi still follow CALC_LATTICE_INVERSE. */
int rsd_per_grpery_inv_fri_images(int n, const double* pos, const double* radii, rsd_lattice* lat, double ews_err,
                                  int nr, int ni, int kmax2, int correction_method, double mu_f, double* app,
                                  double* apq, double* aqq, int* n_overlap);

/* GRPERY_INV_FRI, GRPY_YZD.f90:14-264: the free-space (IsPeriod = false) grand
 * mobility.  Pair separations are taken as they are (no images); overlapping
 * pairs (r <= a_i + a_j) get the Rotne-Prager-Yamakawa blocks (correction_method
 * = 1, YAMAKAWA_IJ, GRPY_YZD.f90:768-786) or the regularised RPY (2) in place of
 * the free one.  Outputs as rsd_per_grpery_inv_fri. */
int rsd_grpery_inv_fri(int n, const double* pos, const double* radii, int correction_method, double mu_f,
                       double* app, double* apq, double* aqq, int* n_overlap);

/* Wall_GRPERY_INV_FRI, Wall_interaction.f90:29-149: the translational mobility
 * TPP(3N,3N) of the particles' images in the wall z = 0 (particles with
 * z <= 10 only), which STEPPER_Stokesian adds to APP(1:3N,1:3N) when
 * Wall_method is set (stepper.f90:67-70). */
int rsd_wall_grpery_inv_fri(int n, const double* pos, const double* radii, double mu_f, double* tpp);

/* lubmxcalc, hydro_lub.f90:22-308 (the reference reads RADII, Nb, IsPeriod,
 * mu_f and LR/LI from modules; here they are arguments).  rfu(6N,6N),
 * rfe(6N,5N), rse(5N,5N) out; *n_contact counts s<=2 reflections. */
int rsd_lubmxcalc(int n, int nb, int is_period, const double* pos, const double* radii, const rsd_lattice* lat,
                  double mu_f, double* rfu, double* rfe, double* rse, int* n_contact);

/* MATREV, matrices.f90:4-34: in-place SPD inverse of a(n,n) using its upper
 * triangle; the result is full (mirrored).  Returns INFO. */
int rsd_matrev(int n, double* a);

/* CHOLESKY, matrices.f90:126-157: the lower Cholesky factor of a(n,n) in
 * place (its lower triangle read), strict upper triangle zeroed.  Returns
 * INFO > 0 for a leading minor that is not positive definite (where the
 * reference stops), with the failed pivot's NaN in the factor. */
int rsd_cholesky(int n, double* a);

/* MobToResist_FTS, matrices.f90:393-416. */
int rsd_mob_to_resist_fts(int n, const double* app, const double* apq, const double* aqq, double* rpp, double* rpq,
                          double* rqq);

/* arrangement, matrices.f90:432-445: mu = RPQ^T inv(RPP) RPQ - RQQ. */
int rsd_arrangement(int n, const double* rpp, const double* rpq, const double* rqq, double* mu);

/* STEPPER_NOGmres, stepper.f90:348-410 (u_brown = 0 and Fe_inter = 0, as
 * STEPPER_Stokesian passes them): U_par(6N), SijN_hyd(5N) and RPP(6N,6N). */
int rsd_stepper_nogmres(int n, const double* app, const double* apq, const double* aqq, const double* rfu,
                        const double* rfe, const double* rse, const double* fe, const double* ein, double* u_par,
                        double* sijn_hyd, double* rpp);

/* ------------------------------------------------ device-resident stepping */
typedef struct rsd_ctx rsd_ctx;

/* Allocates the device state of an n-particle suspension (all the dense
 * 11N and 6N matrices live in HBM for the context's lifetime). */
int rsd_ctx_create(rsd_ctx** ctx, int n, const rsd_step_config* cfg);
int rsd_ctx_destroy(rsd_ctx* ctx);
/* conf(3,N), radii(N), u_pos(6N) (u_pos may be NULL: zero). */
int rsd_ctx_set_state(rsd_ctx* ctx, const double* conf, const double* radii, const double* u_pos);
int rsd_ctx_get_state(rsd_ctx* ctx, double* conf, double* u_pos);
/* STEPPER_Stokesian, stepper.f90:33-156, on the periodic FTS path, with
 * INIT_u_bg (csd_post.f90:442-479) applied first as main.f90:193 does.
 * yeta_mu[5] as yetamu_solve (csd_post.f90:313-367).  The caller advances t.
 * A return > 0 is the INFO of MobToResist_FTS's MATREV of the grand mobility
 * (the order of its first non-positive leading minor) when that failed, else
 * that of the sub-steps' MATREV of RPP; the outputs are then as meaningless
 * as the reference's, which prints the INFO and carries on. */
int rsd_ctx_step(rsd_ctx* ctx, double dt, double t, double yeta_mu[5]);
/* Outputs of the last step: U_par(6N) relative to the background flow,
 * SijN_hyd(5N), and NT_DEM (collision sub-steps). */
int rsd_ctx_get_last(rsd_ctx* ctx, double* u_par, double* sijn_hyd, int* nt_dem);
/* Per-stage device times of the last step in milliseconds (HIP events):
 * [0] Ewald assembly, [1] 11N inverse, [2] lubrication + R assembly,
 * [3] 6N inverse, [4] matvecs + update, [5] whole step. */
int rsd_ctx_stage_times(rsd_ctx* ctx, double ms[6]);

/* Solver for the later collision sub-steps of rsd_ctx_step (no counterpart
 * in the reference, which factors and solves at every sub-step,
 * stepper.f90:95-143 -> STEPPER_NOGmres:348-410):
 *   RSD_SUBSTEP_DIRECT  factor RPP and solve at every sub-step;
 *   RSD_SUBSTEP_PCG     (default) the first sub-step as DIRECT, whose factor
 *                       also gives inv(RPP_1) in FP32; later sub-steps run
 *                       FP64 conjugate gradients on RPP_k preconditioned by it,
 *                       from the last two sub-steps' solutions extrapolated, until
 *                       ||b - RPP_k x|| <= tol ||b|| (default 1e-14), one
 *                       launch per CG phase with a host check every few
 *                       iterations.  If any solve has not converged after
 *                       max_iter iterations (default 60), or finds RPP_k not
 *                       SPD, the step's sub-steps are redone from the step's
 *                       start as DIRECT.
 *   RSD_SUBSTEP_PCG_ONE_LAUNCH  as PCG, each solve one cooperative launch with
 *                       grid barriers and no host check (RPP's order <= 16384
 *                       and a device that runs cooperative launches; as PCG
 *                       otherwise).  Slower than PCG on MI355X at N=512.
 * Steps with NT_DEM = 1 are always solved directly. */
#define RSD_SUBSTEP_DIRECT 0
#define RSD_SUBSTEP_PCG 1
#define RSD_SUBSTEP_PCG_ONE_LAUNCH 2
int rsd_ctx_set_substep_solver(rsd_ctx* ctx, int mode, double tol, int max_iter);
/* How rsd_ctx_step forms the resistance blocks from the grand mobility
 * G = [A B; B^T C] (no counterpart in the reference, whose MobToResist_FTS,
 * matrices.f90:393-416, inverts G whole):
 *   RSD_RESIST_EXPLICIT  W = inv(G) (factor, triangular inverse, product);
 *   RSD_RESIST_SCHUR     (default) W_PP = inv(A - B inv(C) B^T), and W_PQ, W_QQ
 *                        only as the products with vectors the step needs
 *                        (about half the FP64 work).  If any factorisation of
 *                        the step reports INFO > 0, the step is redone with
 *                        EXPLICIT, whose INFO is the reference's. */
/* The geometry rsd_ctx_step runs in (the reference's IsPeriod = simplePeriod
 * and Wall_method flags, stepper.f90:54-70, 160-227; default periodic, no wall):
 *   is_period = 1  Ewald grand mobility under Lees-Edwards images, minimum-
 *                  image lubrication, the Lees-Edwards wrap after each move;
 *   is_period = 0  free-space grand mobility (GRPERY_INV_FRI), lubrication on
 *                  raw separations, no wrap (yeta_mu still divides by the
 *                  box volume LB(1) LB(2) LB(3));
 *   wall_method = 1  the wall images' TT mobility (Wall_GRPERY_INV_FRI) added
 *                  to APP's TT block before the solve. */
int rsd_ctx_set_geometry(rsd_ctx* ctx, int is_period, int wall_method);
/* The Ewald image set rsd_ctx_step sums (rsd_per_grpery_inv_fri_images'
 * nr, ni, kmax2); nr < 0 restores the reference's rule (the default). */
int rsd_ctx_set_ewald_images(rsd_ctx* ctx, int nr, int ni, int kmax2);
/* Brownian motion, the reference's BROWN branch (stepper.f90:72-86): with
 * kbt > 0 each step factors APP(1:3N,1:3N) (CHOLESKY, matrices.f90:126-157;
 * a failure returns RSD_ERR_NOT_SPD, where the reference stops), forms
 * VB = sqrt(2 kBT) A z / DT and the drift DMDX of DMDX_CAL (stepper.f90:650-688,
 * two more grand-mobility assemblies at conf +- 0.0005 z), and
 * u_brown = VB + kBT DMDX.  As in the reference's periodic FTS path, u_brown
 * moves no particle: it enters the Brownian stresslet SijN_B = RPQ^T UB
 * (stepper.f90:407), which yeta_mu(1) includes and yeta_mu(4) reports
 * (csd_post.f90:323-341).  z (3N standard normals) is the one given by
 * rsd_ctx_set_brownian_z for the next step, else drawn on the device by a
 * counter-based generator (Philox4x32-10 keyed by seed, counted by step;
 * Box-Muller's cosine branch, as GAUSS, stepper.f90:638-648).  kbt = 0: off. */
int rsd_ctx_set_brownian(rsd_ctx* ctx, double kbt, unsigned long long seed);
int rsd_ctx_set_brownian_z(rsd_ctx* ctx, const double* z);
/* The last step's z (3N), u_brown (3N) and SijN_B (5N); NULL skips one. */
int rsd_ctx_get_brownian(rsd_ctx* ctx, double* z, double* u_brown, double* sijn_b);
#define RSD_RESIST_EXPLICIT 0
#define RSD_RESIST_SCHUR 1
/*   RSD_RESIST_ITERATIVE  the reference's iterative formulation (STEPPER_Gmres ->
 *                        hydro_SD_FTS, stepper.f90:450-546, which solves with
 *                        GMRES, gmres_mod_gmres.f90:23-61, instead of inverting
 *                        M): each sub-step solves the mobility form
 *                        K x = [Fe; -EIN], K = [(I + R_lub M)_P; M_Q], by
 *                        flexible GMRES in FP64 with products by M (dense) and
 *                        by the lubrication in 11x11 neighbour blocks, then
 *                        U = (M x)_P, SijN_hyd = (x + R_lub M x)_Q.  Tuned by
 *                        rsd_ctx_set_iterative; a solve that does not converge
 *                        redoes the step on the direct Schur route. */
#define RSD_RESIST_ITERATIVE 2
int rsd_ctx_set_resistance_route(rsd_ctx* ctx, int route);
/* Preconditioners of the iterative route (no counterpart in the reference,
 * whose GMRES is unpreconditioned):
 *   RSD_PRECOND_NONE         none, as the reference;
 *   RSD_PRECOND_BLOCK_JACOBI the inverses of K's 11x11 diagonal blocks;
 *   RSD_PRECOND_MIXED        inner_iters FP32 GMRES iterations on FP32 copies
 *                            of M and the lubrication blocks (block-Jacobi
 *                            inside): FP32 bulk work, FP64 outer refinement
 *                            (default);
 *   RSD_PRECOND_LEARNED      block-Jacobi plus the learned pair correction
 *                            of the model attached with rsd_ctx_set_gnn
 *                            (block-Jacobi alone without one);
 *   RSD_PRECOND_MIXED_LEARNED the FP32 inner GMRES of RSD_PRECOND_MIXED with
 *                            the learned preconditioner inside it.
 * tol: relative residual ||b - K x|| <= tol ||b|| (default 1e-12).  The FP64
 * rounding of R_lub (M x), whose entries grow as 1/gap, floors the residual
 * near 1e-12 for dense suspensions: a restart cycle that stalls there within
 * 100 tol is taken as converged.  max_iter (default 400) outer GMRES
 * iterations per solve; inner_iters (default 20, at most 31). */
#define RSD_PRECOND_NONE 0
#define RSD_PRECOND_BLOCK_JACOBI 1
#define RSD_PRECOND_MIXED 2
#define RSD_PRECOND_LEARNED 3
#define RSD_PRECOND_MIXED_LEARNED 4
int rsd_ctx_set_iterative(rsd_ctx* ctx, int precond, double tol, int max_iter, int inner_iters);
/* on = 1 (the default): the mixed preconditioners' FP32 inner GMRES is captured
 * once as a HIP graph and replayed (one host launch per inner solve in place
 * of about 13 per inner iteration); 0: launched kernel by kernel.  The two
 * perform the same arithmetic. */
int rsd_ctx_set_iterative_graphs(rsd_ctx* ctx, int on);
/* The last step's total GMRES iterations, its largest final relative residual,
 * and 1 if a solve did not converge and the step was redone directly. */
int rsd_ctx_iterative_stats(rsd_ctx* ctx, int* iterations, double* relres, int* fell_back);

/* The learned preconditioner (no counterpart in the reference).  A model is a
 * graph network over the lubrication's neighbour pairs (3 message-passing
 * layers, hidden width 64, FP32, RSD_GNN_PARAMS parameters) that predicts,
 * for each pair (g, d), a weight w in (0, 1) for the pieces of the inverse of
 * K's 22x22 pair block added to block-Jacobi: P_gg = D_g + sum_d w (A_gd -
 * D_g), P_gd = w B_gd (a weighted additive Schwarz over pairs).
 * rsd_gnn_create initialises it (seeded) with zero output weights: every pair
 * then weighs 1/2.  Parameters travel as float arrays. */
typedef struct rsd_gnn rsd_gnn;
#define RSD_GNN_PARAMS 50629
int rsd_gnn_create(rsd_gnn** gnn, unsigned long long seed);
int rsd_gnn_destroy(rsd_gnn* gnn);
int rsd_gnn_get_params(rsd_gnn* gnn, float* params);
int rsd_gnn_set_params(rsd_gnn* gnn, const float* params);
/* Attach a model to a context for RSD_PRECOND_LEARNED / _MIXED_LEARNED (NULL
 * detaches; the model must outlive its use). */
int rsd_ctx_set_gnn(rsd_ctx* ctx, rsd_gnn* gnn);
/* Assemble the iterative route's system at the context's current positions
 * and time t without stepping: the grand mobility, the block-sparse
 * lubrication, block-Jacobi and the learned preconditioner's pair pieces. */
int rsd_ctx_prepare_iterative(rsd_ctx* ctx, double t);
/* The prepared system on the host (each pointer may be NULL): M (11N x 11N,
 * column-major), the lubrication R (11N x 11N, from its blocks), nnb[N],
 * col[N cap] and dist[N cap] (neighbour d and pair distance of slot q of
 * particle g at g cap + q), and cap. */
int rsd_ctx_iterative_system(rsd_ctx* ctx, double* M, double* R, int* nnb, int* col, double* dist, int* cap);
/* The training loss sum_z |K P z - z|^2 / sum_z |z|^2 of the prepared system
 * for nz Gaussian vectors z drawn from seed (z out: nz x 11N when not NULL)
 * and, when grad is not NULL, its gradient with respect to the parameters. */
int rsd_ctx_gnn_loss(rsd_ctx* ctx, rsd_gnn* gnn, int nz, unsigned long long seed, double* loss, float* grad,
                     double* z);
/* iters Adam steps (learning rate lr) of the model on the prepared system, each
 * on nz fresh Gaussian vectors, minimising the log of the loss; losses[iters]
 * (may be NULL) receives each step's loss before its update. */
int rsd_ctx_gnn_train(rsd_ctx* ctx, rsd_gnn* gnn, int iters, int nz, float lr, unsigned long long seed,
                      double* losses);
/* One solve of the prepared system with right-hand side [0; -EIN] from zero
 * with the given preconditioner: GMRES iterations, final relative residual,
 * and the solve's time in ms (the preconditioner's set-up included). */
int rsd_ctx_iterative_solve(rsd_ctx* ctx, int precond, double tol, int max_iter, int* iterations, double* relres,
                            double* ms);
/* The last step's total CG iterations and count of CG solves that did not
 * converge (nonzero: its sub-steps were redone directly). */
int rsd_ctx_substep_stats(rsd_ctx* ctx, int* cg_iterations, int* fallbacks);
/* As rsd_ctx_substep_stats, plus whether the last step was redone through the
 * explicit 11N inverse because a Schur-route factorisation reported INFO > 0
 * (explicit_redo = 1: about twice the dense work; the step's INFO is then the
 * explicit route's, which MATREV would report, matrices.f90:19-24). */
int rsd_ctx_step_stats(rsd_ctx* ctx, int* cg_iterations, int* cg_fallbacks, int* explicit_redo);

/* Batches (no counterpart in the reference, which runs one suspension per
 * process): independent suspensions, each its own context (its own streams,
 * sizes and settings), stepped together.  rsd_batch_step runs `width` host
 * threads (default min(count, 8)) that take the contexts, largest N first,
 * from a shared counter, so that the steps of small suspensions overlap on
 * the GPU; every step's arithmetic is its context's own, so a batch gives the
 * numbers each context gives alone.  t[count] (NULL: all 0), yeta_mu[5 count]
 * and info[count] (INFO of each step) may be NULL.  Kernel timing must be off.
 * rsd_batch_autotune times one batch step per candidate width (after one
 * warm-up step, each from the same saved states, which it restores at the
 * end) and keeps the fastest; ms[ncand] receives the times.  The contexts
 * must outlive the batch. */
typedef struct rsd_batch rsd_batch;
int rsd_batch_create(rsd_batch** batch, int count, rsd_ctx* const* ctxs);
int rsd_batch_destroy(rsd_batch* batch);
int rsd_batch_set_width(rsd_batch* batch, int width);
int rsd_batch_get_width(rsd_batch* batch, int* width);
int rsd_batch_step(rsd_batch* batch, double dt, const double* t, double* yeta_mu, int* info);
int rsd_batch_autotune(rsd_batch* batch, double dt, const double* t, const int* cand, int ncand, int* best,
                       double* ms);

/* Kernel timing (no counterpart in the reference, whose only timers are
 * cpu_time tic/toc in csd_post.f90:188-238): when enabled, each launch of a
 * kernel family is bracketed by HIP events on its own stream.  Families:
 * 0 gemm64, 1 potrf_dag (dataflow Cholesky), 2 potrs (triangular solves), 3 ewald_pair, 4 ewald_self,
 * 5 lub, 6 gemv, 7 pcg (one conjugate-gradient sub-step solve per record), 8 gemv_split_f64 and
 * 9 gemv_split_f32 (the iterative route's products with M, FP64 and FP32), 10 inner_gmres_graph
 * (one replay of the mixed preconditioner's captured FP32 GMRES; the launches inside a graph are
 * timed only as the graph).  flops/bytes are the algorithmic work of the launches. */
int rsd_profile_enable(int on);
/* As rsd_profile_enable, for the families whose bits are set in mask (bit f
 * for family f): events around fewer launches cost the host less. */
int rsd_profile_enable_families(unsigned mask);
int rsd_profile_read(int family, long* launches, double* ms, double* flops, double* bytes);
const char* rsd_profile_family_name(int family);

/* Host-array convenience form of STEPPER_Stokesian (one step). */
int rsd_stepper_stokesian(int n, double* conf, const double* radii, double dt, double t, const rsd_step_config* cfg,
                          double* yeta_mu, double* u_pos);

/* ------------------------------------------------- on-disk formats, units */
/* control_file.dat, read positionally as INIT_READ_CONTROL does
 * (initial_conf.f90:109-189): one field (or vector) per READ, in this order.
 * Logicals are 0/1.  kbt is read only when brown is set. */
typedef struct rsd_control {
  int correction_method, fts_method, filament_solve_implicit, uselub, use_dem, usecollision, use_dem_stepper,
      use_gmres, xshear, is_period, simple_period, oscillation_shear, fix, is_case_lattice, beginstep, interval,
      num_threads, timestep;
  double radii_test, radii_true, g_gravity, u_bg[3], omega_bg[3], ei_bg[6], mu_f, pho_f, re_number, gamma_alpha,
      youngs_module, poisson_ratio, ksp, erest, pho_par, linerr;
  int collision_condition, mu_calculate;
  double t0, end_time, base_dt;
  int k_write;
  double lj_eps, lj_sigma, lj_cut;
  int boundary, wall_method, use_dlvo;
  double debye_length, z0, lambda_l, permittivity, potential, ah, surfaceenergy;
  int brown;
  double kbt;
} rsd_control;

/* The step's inputs in the reference's nondimensional units:
 * dimentionless_calc / dimentionless_translate (mymodules.f90:808-881), and
 * frequency = 1/lambda_time, GAMMA = Gamma_alpha * frequency (main.f90:148-155). */
typedef struct rsd_units {
  double lambda_time, mu_f, gravity, end_time, base_dt, t0, dt_dem0, kbt, frequency, gamma;
} rsd_units;

/* INIT_READ_CONTROL, initial_conf.f90:109-189. */
int rsd_read_control(const char* path, rsd_control* ctl);
/* INITIAL_SIZE + INITIAL_CONF for IsCaseLattice = 0 (initial_conf.f90:356-476):
 * *n = Np and lb, hot_start always; with conf != NULL (capacity cap particles)
 * also t (the last row's T), conf(3,Np), radii(Np) and, when hot_start, the
 * translational u_pos(3Np) of the rows T,x,y,z,a,Ux,Uy,Uz that VMD_WRITE writes.
 * Rigid conglomerates and filaments return RSD_ERR_UNSUPPORTED. */
int rsd_read_initial_config(const char* path, int cap, int* n, double lb[3], double* t, double* conf, double* radii,
                            double* u_pos, int* hot_start);
int rsd_nondimensionalise(const rsd_control* ctl, rsd_units* units);
/* VMD_WRITE, initial_conf.f90:250-262: one row per free particle,
 * T,x,y,z,a,Ux,Uy,Uz in (13e24.15e3); append = 0 truncates the file. */
int rsd_write_vtf_frame(const char* path, int append, int n, int nb, const double* conf, double t,
                        const double* radii, const double* u_pos);
/* main.f90:197-201: T, yeta_mu(1:5), LB(1), frequency, GAMMA, GAMMAangle in (12ES24.15). */
int rsd_write_yeta_mu(const char* path, int append, double t, const double yeta_mu[5], double lb1, double frequency,
                      double gamma, double gamma_angle);

#ifdef __cplusplus
}
#endif
#endif /* RSD_H_ */
