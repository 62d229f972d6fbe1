/*
 * pdm.h — C ABI of libpdm, the MI355X-native projective-dynamics engine for
 * diff_fem's corotated-tetrahedral FEM (forward step and its gradient).
 *
 * Plain C types only; every function returns 0 on success and a non-zero
 * status otherwise, with a message in pdm_last_error(). No exceptions cross
 * this boundary. All state vectors are in the caller's layout: per rollout the
 * reference's AoS x[3*i + c] (python/pydifem.cc passes std::vector<double>),
 * rollouts stacked: x[r * 3N + 3*i + c].
 *
 * Entry points and the reference interface each replaces:
 *   pdm_scene_create   <- Environment(total_step, config) + Dfobj::SetDfobj /
 *                         Dfobj::Initialize + World::Initialize
 *                         (sim/Environment.h:85-225, sim/Dfobj.h:142-314,
 *                         442-611, sim/fem/World.h:248-384)
 *   pdm_set_material   <- forward_all's mMu update + World::PreComputation
 *                         (python/pydifem.cc:207-221, sim/fem/World.h:1089-1139)
 *   pdm_forward(_host) <- forward_all / forward_act: Environment::Step x steps
 *                         (python/pydifem.cc:38-86, 182-259;
 *                         sim/Environment.h:228-293; sim/fem/World.h:673-1029)
 *   pdm_backward(_host)<- backward_all / backward_act: CppAD Jacobian of
 *                         <dldq,q_T>+<dldqd,qd_T> (python/pydifem.cc:90-180, 261-301)
 *   pdm_solve_host     <- SimplicialLDLT::solve on H2ML (sim/fem/World.h:1660)
 *   pdm_force_host     <- World::EvaluateDVector + J*d (sim/fem/World.h:1194-1210, 901)
 *   pdm_host_analyse / pdm_host_solve <- World::PreComputation + FactorizeLDLT on
 *                         the CPU (sim/fem/World.h:1089-1139, 1212-1232)
 */
#ifndef PDM_H_
#define PDM_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdm_scene_s* pdm_scene;

/* Scene description (sim/Environment.h:109-210 keys; sim/Dfobj.h:142-314). */
typedef struct pdm_scene_desc {
  int n_vertices;            /* N */
  const double* x_rest;      /* 3N, rest positions (already scaled) */
  int n_tets;                /* T */
  const int* tets;           /* 4T, 0-based, any orientation */
  int mass_same;             /* "Is Mass Same": volume-lumped (1) or uniform (0) */
  double total_mass;         /* "Mass" */
  const double* E;           /* T Young's moduli (NULL: 2e5) */
  double nu_fwd;             /* nu entering mMu = E/(1+nu) */
  const double* nu_ctor;     /* T nu of the volume-term weight (NULL: 0.4) */
  int n_attach;              /* "Attach Verts" */
  const int* attach_idx;
  double attach_w;           /* "Attach Stiffness" */
  int n_segments;            /* muscle segments (50 per fibre in the reference) */
  const double* seg_start;   /* 3 per segment */
  const double* seg_end;     /* 3 per segment */
  const int* seg_group;      /* activation group of each segment */
  int n_groups;              /* activation groups (columns of the act arrays) */
  double k_muscle;           /* "Muscle Stiffness" */
  double h;                  /* 1 / "Simulation HZ" */
  int max_iter;              /* "Max Iteration" (PD loop) */
  double damping;            /* "Damping coefficient" */
  double gravity[3];         /* "Gravity" */
  int gravity_on;            /* "Enable Gravity" */
  int contact;               /* "Handle collision" */
  int max_iter1;             /* contact Jacobi cap ("Max Iteration") */
  double alpha;              /* "World Alpha" */
  double mu;                 /* friction coefficient of obstacle contacts */
  int n_planes;              /* static planes: point(3), normal(3) */
  const double* planes;
  int n_spheres;             /* static spheres: centre(3), radius */
  const double* spheres;
  /* Several bodies in one world: one Dfobj per mesh (Environment.h:109-210,
   * the "Meshes" list), each with its own "Mass" lumped on its own volume
   * (Dfobj.h:506-525) and its own Poisson ratio in mMu = E/(1+nu)
   * (CorotateFEMConstraint.h:62-77). The PD loop, its early exit and the
   * contact phase stay those of the one World (World.h:835-1023). */
  const double* nu_fwd_tet;  /* T values of nu entering mMu (NULL: nu_fwd for every tet) */
  int n_bodies;              /* 0: one body of total_mass */
  const int* body_of_tet;    /* T body ids in [0, n_bodies) */
  const double* body_mass;   /* n_bodies masses */
  /* "Is underwater" (Environment.h:162-168, 236-250): no gravity; each step
   * starts with v* = v + h M^-1 f, f the drag and lift of every surface face
   * at the step-start state (Dfobj::ComputeDragForces, Dfobj.h:846-994). */
  int underwater;
  /* Static obstacle triangles ("Obstacles" meshes, Dfobstacle.h:110-160):
   * 9 values per triangle (three corners). Surface vertices collide with
   * them by ArcSim's vertex-face test (collision.cpp:127-278). */
  int n_tris;
  const double* tris;
  /* "Self-collision" (World.h:1246-1315, 1446-1600): surface vertices against
   * the body's own surface faces, with "World Friction" self_mu. The reverse
   * sweep holds group membership, stick/slide flags and impact normals as
   * executed (the reference computes them from plain doubles). */
  int self_collision;
  double self_mu;
} pdm_scene_desc;

typedef struct pdm_info {
  int n_vertices, n_tets, n_muscles, n_groups;
  int n_supernodes, n_levels, max_front, max_width;
  long long nnz_factor;      /* stored entries of the supernodal factor L */
  long long nnz_lower;       /* entries of the lower-trapezoidal supernode blocks (algorithmic nnz(L)) */
  double factor_flops;       /* flops of one numeric factorisation */
  double reg_used;           /* regularisation applied by the ladder (0 if none) */
  int n_contact_vertices;
  int batch;                 /* rollouts R */
  int precision;             /* PDM_FP64, PDM_MIXED or PDM_MIXED_ADJOINT */
  int n_drag_faces;          /* surface faces carrying drag (0 unless underwater) */
  int n_obstacle_tris;       /* static obstacle triangles */
  int n_self_faces;          /* surface faces tested for self contact */
  int top_rows;              /* rows of the dense top-separator blocks (padded; 0: none) */
  int top_blocks;            /* elimination trees with a dense top-separator block */
  int leaf_blocks;           /* leaf subtrees solved as dense products */
} pdm_info;

/* Arithmetic of pdm_set_batch:
 *   PDM_FP64           FP64 everywhere;
 *   PDM_MIXED          FP32 local step (forward and its VJP in the reverse
 *                      sweep), FP64 factor, solves, contact and accumulation;
 *   PDM_MIXED_ADJOINT  FP64 forward (the trajectory, iteration counts and
 *                      contact branches are the FP64 path's), FP32 local-step
 *                      VJP in the reverse sweep, FP64 elsewhere. */
enum { PDM_FP64 = 0, PDM_MIXED = 1, PDM_MIXED_ADJOINT = 2 };

/* Flags of pdm_set_batch. */
enum {
  PDM_HOST_EARLY_EXIT = 1, /* read the convergence test back each iteration and
                              stop launching (the reference's per-iteration break);
                              without it converged rollouts are masked on device */
  PDM_USE_GRAPH = 2,       /* forward (and the adjoint's recompute): capture the step
                              start + PD iterations once per call as a hipGraph (one for
                              taped steps, whose tape slot is read from a device
                              counter, one for untaped steps) and replay it every step,
                              with the early exit decided on the device (converged
                              rollouts masked, gather and solves skipped once none
                              iterates); the contact phase and the adjoint's reverse
                              sweep stay eager */
  PDM_TIME_SOLVES = 4,     /* bracket every triangular solve with HIP events on the
                              library stream (read with pdm_get_solve_timing) */
  PDM_NO_AUTOTUNE = 8      /* keep the default solve launch shapes instead of timing
                              the candidates once per right-hand-side count */
};

const char* pdm_last_error(void);
int pdm_device_count(int* n);

int pdm_scene_create(const pdm_scene_desc* desc, pdm_scene* out);
void pdm_scene_destroy(pdm_scene s);
int pdm_scene_info(pdm_scene s, pdm_info* info);

/* Bind a HIP stream (hipStream_t as void*; NULL = the library's own stream). */
int pdm_set_stream(pdm_scene s, void* stream);

/* Material update + numeric re-factorisation on the device. E: T values
 * (caller tet order) or NULL to keep; nu_fwd as forward_all's tau[1];
 * nu_fwd = NaN keeps the current per-tet nu (from the scene description). */
int pdm_set_material(pdm_scene s, const double* E, double nu_fwd);

/* Numeric re-factorisation of K_s = M/h^2 + L for the current material, with
 * the regularisation ladder of World::FactorizeLDLT (World.h:1212-1232;
 * replaces World::PreComputation's refactor, World.h:1089-1139, that
 * forward_all runs on every call, pydifem.cc:207-221). on_host = 0: on the
 * device, as pdm_scene_create and pdm_set_material do; 1: the same
 * factorisation in host code, uploaded (a cross-check of the device path).
 * The time lands in pdm_get_timing's factor_ms. */
int pdm_refactor(pdm_scene s, int on_host);

/* Copy the solve operand W = [inv(L_ss); -L_below inv(L_ss)] of every
 * supernode (column-major m x w blocks in factor order; n = nnz_factor).
 * No counterpart in the reference (a diagnostic of the factorisation). */
int pdm_get_factor_w(pdm_scene s, double* W, long long n);

/* Rollout batch R, arithmetic precision (PDM_FP64 everywhere, or PDM_MIXED:
 * FP32 local step, FP64 factor/solves/accumulation), flags, and the
 * checkpoint interval used by recorded forwards (steps between stored states). */
int pdm_set_batch(pdm_scene s, int R, int precision, int flags, int checkpoint_every);

/* Forward rollout of `steps` steps for all R rollouts. Device pointers:
 * x0, v0, xT, vT: R*3N doubles; act: R*steps*n_groups doubles or NULL (zero
 * activation). record != 0 keeps checkpoints for pdm_backward. */
int pdm_forward(pdm_scene s, int steps, const double* x0, const double* v0, const double* act, double* xT, double* vT,
                int record);
int pdm_forward_host(pdm_scene s, int steps, const double* x0, const double* v0, const double* act, double* xT,
                     double* vT, int record);

/* Reverse sweep over the last recorded forward: gradient of
 * sum_r <gxT_r, x_T,r> + <gvT_r, v_T,r> w.r.t. x0, v0 (R*3N), per-tet E (T,
 * caller order, summed over rollouts), nu_fwd (scalar) and act
 * (R*steps*n_groups). Device pointers; any output may be NULL. */
int pdm_backward(pdm_scene s, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gE, double* gnu,
                 double* gact);
int pdm_backward_host(pdm_scene s, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gE,
                      double* gnu, double* gact);

/* Executed PD iterations of the last forward: R*steps ints (host). */
int pdm_get_iterations(pdm_scene s, int* iters);

/* Contact phase of the last forward: detection passes that found impacts and
 * executed Jacobi iterations, R*steps ints each (host). */
int pdm_get_contact_stats(pdm_scene s, int* passes, int* jacobi_iters);

/* Diagnostics on host arrays (caller layout, R rollouts): solve
 * (M/h^2 + L) x = b with the device factor; evaluate J*d at positions x. */
int pdm_solve_host(pdm_scene s, const double* b, double* x);
int pdm_force_host(pdm_scene s, const double* x, const double* act_step, double* Jd);

/* Host-only diagnostics (no device needed): run scene setup, ordering,
 * symbolic analysis and the numeric factorisation on the CPU; report the
 * factor statistics, or solve (M/h^2 + L) x = b for R rollouts with the same
 * supernodal schedule the device kernels use. */
int pdm_host_analyse(const pdm_scene_desc* desc, pdm_info* info);
int pdm_host_solve(const pdm_scene_desc* desc, int R, const double* b, double* x);

/* Time `reps` back-to-back triangular solves (forward + backward sweep over
 * all R*3 right-hand sides currently in the engine's rhs buffer) with HIP
 * events on the library stream; returns the mean ms per solve. */
int pdm_time_solve(pdm_scene s, int reps, double* ms_per_solve);

/* With PDM_TIME_SOLVES: number of triangular solves (forward + backward sweep
 * over all 3R right-hand sides) run since the last call and their summed
 * duration in ms (HIP events around each solve on the library stream); the
 * counters restart at zero. No counterpart in the reference (measurement). */
int pdm_get_solve_timing(pdm_scene s, long long* n_solves, double* total_ms);

/* Launch shapes of the level kernels for the current batch (chosen by the
 * autotuner, or the defaults under PDM_NO_AUTOTUNE): for level l < cap,
 * column tile ct_*[l] and threads per block thr_*[l] of the forward (fwd) and
 * backward (bwd) sweep. *n_levels receives the level count, *n_chains the
 * number of column chains the solve is split into (each on its own stream),
 * *subtree_cut the level below which whole subtrees run in one launch (0: none).
 * No counterpart in the reference (launch tuning). */
int pdm_get_solve_shapes(pdm_scene s, int cap, int* n_levels, int* ct_fwd, int* thr_fwd, int* ct_bwd, int* thr_bwd,
                         int* n_chains, int* subtree_cut);

/* Override the autotuner's whole-solve plan for the current batch: levels below
 * subtree_cut run as whole subtrees (0: none), and the columns are split into
 * n_chains chains on separate streams (1..4). Results do not depend on the
 * plan. No counterpart in the reference (launch tuning). */
int pdm_set_solve_plan(pdm_scene s, int subtree_cut, int n_chains);

/* Dense top-separator blocks of the solve: in every elimination tree, the
 * supernodes of the top heights that own at most max_rows rows together (the
 * largest such set) are solved as one dense product x_t = Z f_t with
 * Z = (L_t L_t^T)^{-1} formed on the device at every factorisation, in place
 * of their levels in both sweeps (the other supernodes run as before and see
 * the same values up to rounding). max_rows = 0 turns the blocks off; the
 * default is 1536. Re-plans the solve (launch shapes are re-chosen on the next
 * call). No counterpart in the reference (it solves with Eigen's LDL^T). */
int pdm_set_dense_top(pdm_scene s, int max_rows);

/* Dense leaf subtrees of the solve: below the top-separator blocks, every
 * maximal subtree whose supernodes own at most max_rows rows together is
 * solved as two dense products, z = K_s^{-1} b_s with its root's update block
 * u = G b_s before the sweeps and x_s = z + G^T x_above after them (K_s^{-1}
 * and G formed on the device at every factorisation). max_rows = 0 turns them
 * off; the default is 512. Re-plans the solve. No counterpart in the
 * reference. */
int pdm_set_dense_leaves(pdm_scene s, int max_rows);

/* Timing of the last forward/backward (ms, HIP events on the library stream). */
int pdm_get_timing(pdm_scene s, double* fwd_ms, double* bwd_ms, double* factor_ms);

/* Step-block graph replays (PDM_USE_GRAPH) since the scene was created, taped
 * and untaped steps together. No counterpart in the reference (measurement). */
int pdm_get_graph_launches(pdm_scene s, long long* n);

/* Largest device scratch pdm_solve_noalloc uses: the engine-layout right-hand
 * sides (N x 3R doubles) plus the multifrontal update blocks (U_rows x 3R
 * doubles), each rounded up to 256 B. */
#define PDM_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the solve of pdm_solve_host, (M/h^2 + L) x = b for R
 * rollouts, on device buffers b and x (R*3N doubles each, caller layout; they
 * may alias), enqueued on the scene's stream without waiting for it. Its
 * working memory is one fixed array per GPU of PDM_NOALLOC_SCRATCH_BYTES,
 * held until the process ends; R is free of the scene's batch but must keep
 * (N + U_rows) * 3R * 8 B within that size (else an error). It uses the
 * default solve launch shapes and the level sweeps over all supernodes,
 * without the dense top-separator blocks, so it agrees with pdm_solve_host to
 * rounding.
 * Every call on one GPU shares that array, so calls
 * that overlap in time on one GPU (other streams, other threads) race; the
 * scene's device must be current. */
int pdm_solve_noalloc(pdm_scene s, int R, const double* b, double* x);

#ifdef __cplusplus
}
#endif
#endif /* PDM_H_ */
