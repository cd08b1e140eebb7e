/*
 * pgt_ik.h -- C ABI of the MI355X Irving-Kirkwood stress-field engine
 * (libpgt_ik.so).
 *
 * Plain C types only: host pointers unless a name says `d_` (device pointer
 * from hipMalloc or any HIP allocator on the engine's device).  Every entry
 * point returns PGT_OK (0) or a negative error code; pgt_last_error() returns
 * the calling thread's message for the last failure.  No C++ exception crosses
 * this boundary.  One engine is used by one host thread at a time.
 *
 * The reference (mandadapu-group/pyglasstools) binds these operations through
 * pybind11 classes; each declaration names the reference interface it
 * replaces (paths relative to the reference checkout).
 */
#ifndef PGT_IK_H
#define PGT_IK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGT_OK 0
#define PGT_ERR_ARG (-1)       /* invalid argument (reference: std::invalid_argument -> ValueError) */
#define PGT_ERR_HIP (-2)       /* HIP runtime error, including "no GPU" */
#define PGT_ERR_STATE (-3)     /* call order: e.g. compute before set_particles */
#define PGT_ERR_RANGE (-4)     /* search radius >= half a periodic box length */
#define PGT_ERR_UNSUPPORTED (-5)

/* Pair potentials of the polydisperse m-n + even-polynomial family.
 * Reference: pyglasstools/potential/module-potential.cpp:17-19,26-30 */
enum pgt_potential_kind {
    PGT_POT_POLY12 = 0,  /* Polydisperse12.h:38-54, params {v0, eps, c0, c1, c2} */
    PGT_POT_POLY18 = 1,  /* Polydisperse18.h:37-53, params {v0, eps, c0, c1, c2} */
    PGT_POT_POLY10 = 2,  /* Polydisperse10.h:37-52, params {v0, eps, c0, c1, c2, c3} */
    PGT_POT_POLY106 = 3, /* Polydisperse106.h:37-53, params {v0, eps, c0, c1, c2} */
    PGT_POT_POLYLJ = 4   /* PolydisperseLJ.h:37-53, params {v0, eps, c0, c1, c2} */
};

/* Coarse-graining kernels.  Reference: irvingkirkwood/cgfunc/module-cgfunc.cpp:8-16 */
enum pgt_cg_kind {
    PGT_CG_OCTIC = 0,    /* CGFuncOctic.h:16-29 */
    PGT_CG_RECT = 1,     /* CGFuncRect.h:17-28 */
    PGT_CG_MOLLIFIER = 2 /* CGFuncMollifier.h:11-31 */
};

/* Bond line-integral rule.  Reference: CoarseGrainFunction.h:56-134 */
enum pgt_quad_mode {
    PGT_QUAD_FIXED = 0,   /* FixedPointCGFunc: n-point Gauss-Legendre on [0,1] */
    PGT_QUAD_ADAPTIVE = 1 /* AdaptiveCGFunc: QAG with 15-point Gauss-Kronrod */
};

/* Arithmetic of the field kernels (the reference has FP64 only). */
enum pgt_precision {
    PGT_PREC_FP64 = 0, /* the drop-in: FP64 everywhere */
    PGT_PREC_FP32 = 1  /* FP32 pair forces and kernel weights; FP64 minimum image
                          and FP64 per-grid-point accumulation */
};

/* Field selection bits for pgt_ik_compute*.  Reference field classes:
 * irvingkirkwood/module-irvingkirkwood.cpp:8-13 */
enum pgt_field_bits {
    PGT_FIELD_VIRIAL = 1,  /* VirialStressField2D/3D        (ListofCommonObservables.h:43-84) */
    PGT_FIELD_KINETIC = 2, /* KineticStressField2D/3D       (ListofCommonObservables.h:125-149) */
    PGT_FIELD_ELASTIC = 4  /* ElasticVirialStressField2D/3D (ListofCommonObservables.h:86-123) */
};

/* Everything IrvingKirkwood's constructor receives, flattened:
 * pair potential (PairPotential.h:30-33), coarse-graining function
 * (CoarseGrainFunction.h:64-67,102-105), and Calculator::max_rcut
 * (Calculator.h:37-46), which the caller computes from the diameters of the
 * system the calculator is built with and which stays fixed afterwards. */
typedef struct pgt_ik_params {
    int dim;                  /* 2 or 3: size of the output tensors */
    int potential_kind;       /* enum pgt_potential_kind */
    double scaled_rcut;       /* PairPotential::scaled_rcut */
    double pot_params[6];     /* PairPotential::params */
    int n_pot_params;
    int cg_kind;              /* enum pgt_cg_kind */
    double cg_rcut;           /* CoarseGrainFunction::cg_rcut */
    int quad_mode;            /* enum pgt_quad_mode */
    int quad_order;           /* fixed rule: number of nodes (1..1024) */
    int adapt_limit;          /* adaptive rule: GSL workspace limit (neval), honoured in full: integrals
                                 needing more than 128 intervals get tables of this limit in global
                                 memory (compute and compute_frames; the enqueue and no-allocation
                                 entry points return PGT_ERR_RANGE instead) */
    double adapt_epsrel;
    double adapt_epsabs;
    int precision;            /* enum pgt_precision */
    double max_rcut;          /* j search radius */
} pgt_ik_params;

typedef struct pgt_ik_stats {
    int64_t n_particles;
    int64_t n_grid;
    int64_t n_ball;        /* (grid point, particle in ball) pairs, last compute */
    int64_t n_contrib;     /* U: (grid point, particle, neighbour) with F != 0 */
    int32_t bond_capacity; /* per-particle bond slots in use */
    int32_t max_bonds;     /* largest per-particle bond count seen */
    double gather_ms;      /* HIP-event time of the field gather kernel(s), last compute */
    double total_ms;       /* HIP-event time of the whole last compute */
    int64_t n_pairs;       /* surrogate form: (grid point, bond) pairs of the last compute */
    int64_t n_network;     /* surrogate form: pairs whose bond integral came from the network */
    double surrogate_ms;   /* surrogate form: HIP-event time from the pair list to the field */
} pgt_ik_stats;

typedef struct pgt_ik_engine pgt_ik_engine;

/* Thread-local message for the last failed call on this thread. */
const char *pgt_last_error(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int pgt_device_count(void);

/* Thresholds of the FP32 distance pre-test of the 3D candidate scans for a
 * search radius r, as the engine sets them: coordinates (box and grid) at
 * most max_coord in magnitude, box lengths max_box and min_box over the
 * searched axes, periodic = every searched axis periodic.  FP32 |a|^2 below
 * *in32 is inside r in FP64 and above *out32 outside; without a bound
 * *in32 = -1 and *out32 = inf.  Host code only (no device needed); for
 * checking the bound.  No counterpart in the reference. */
int pgt_pretest_band(double r, double max_coord, double max_box, double min_box, int dim, int periodic,
                     float *in32, float *out32);

/* Create an engine on the current HIP device.
 * Replaces: IrvingKirkwood(ParticleSystem, PairPotential, CoarseGrainFunction,
 * Communicator), irvingkirkwood/IrvingKirkwood.h:12-18, 116. */
int pgt_ik_create(const pgt_ik_params *params, pgt_ik_engine **out);
void pgt_ik_destroy(pgt_ik_engine *e);

/* Box lengths, lower bounds and periodic flags (z non-periodic in 2D).
 * Replaces: SimBox(boxsize, origin, dim) + setBounds, SimBox.cc:3-52. */
int pgt_ik_set_box(pgt_ik_engine *e, const double boxsize[3], const double lowerbound[3],
                   const int periodic[3]);

/* Host particle data, N x 3 row-major positions/velocities/displacements.
 * mass, vel and disp may be NULL (mass 1, zero velocity/displacement).
 * Replaces: ParticleSystem(simbox, N, diameter, mass, position, velocity)
 * (ParticleSystem.cc:3-30) and IrvingKirkwood::setSystemData
 * (Calculator.h:50-53); the neighbour index is built on the GPU at compute. */
int pgt_ik_set_particles(pgt_ik_engine *e, int64_t n, const double *pos, const double *diam,
                         const double *mass, const double *vel, const double *disp);

/* Host grid points, G x 3 row-major.
 * Replaces: the gridpoints argument of IrvingKirkwood::compute (IrvingKirkwood.h:21). */
int pgt_ik_set_grid(pgt_ik_engine *e, int64_t g, const double *grid);

/* Compute the selected fields for the current particles and grid and copy
 * them to host arrays of G x D x D doubles (flat index a + D*b per grid
 * point, the layout gettostring reads, CoarseGrainedField.h:160-183).
 * Unselected outputs may be NULL.
 * Replaces: IrvingKirkwood::compute, irvingkirkwood/IrvingKirkwood.h:74-110. */
int pgt_ik_compute(pgt_ik_engine *e, unsigned fields, double *virial, double *kinetic,
                   double *elastic);

/* Batched, device-resident: nframes systems of n particles each, all on the
 * device, over the grid given by pgt_ik_set_grid.  Frame f reads
 * d_pos + f*3n, d_diam + f*n, d_mass + f*n (NULL: 1), d_vel + f*3n (NULL: 0),
 * d_disp + f*3n (NULL: 0) and writes G*D*D doubles at d_out_* + f*G*D*D.
 * No host synchronisation inside unless a bond list outgrows its capacity.
 * Replaces: the per-frame loop of pyglasstools.analyze (__init__.py:119-122)
 * around IrvingKirkwood::compute, for callers that keep frames on the GPU. */
int pgt_ik_compute_frames(pgt_ik_engine *e, int nframes, int64_t n, const double *d_pos,
                          const double *d_diam, const double *d_mass, const double *d_vel,
                          const double *d_disp, unsigned fields, double *d_out_virial,
                          double *d_out_kinetic, double *d_out_elastic);

/* pgt_ik_compute_frames that allocates nothing.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and statistics as pgt_ik_compute_frames, but every
 * work array (cell list, bond list, statistics) lives in a static device
 * array of 64 MiB per GPU, held until the process ends; the call makes no
 * hipMalloc, hipFree or other allocating call.  Frames are run in launches of
 * at most pgt_ik_set_tuning's frames-per-launch that fit the array; a job
 * whose single frame needs more work memory than that returns PGT_ERR_RANGE
 * (2D FP64 virial + kinetic: about 450 bytes per particle, so up to ~140,000
 * particles per frame).  The engine itself (pgt_ik_create, pgt_ik_set_grid)
 * allocates at set-up as before.  All engines on one GPU share its array and
 * it keeps the last call's contents: calls that overlap on one GPU race, so
 * serialise them.  Synchronises the engine's stream. */
int pgt_ik_compute_frames_noalloc(pgt_ik_engine *e, int nframes, int64_t n, const double *d_pos,
                                  const double *d_diam, const double *d_mass, const double *d_vel,
                                  const double *d_disp, unsigned fields, double *d_out_virial,
                                  double *d_out_kinetic, double *d_out_elastic);

/* Enqueue a job like pgt_ik_compute_frames and return at once: no host
 * synchronisation, no device allocation, no memset, no events, so a caller
 * can capture the launches into a HIP graph (hipStreamBeginCapture on the
 * engine's stream, see pgt_ik_set_stream) and replay them.  Needs an earlier
 * pgt_ik_compute_frames on this engine with the same frame count, particle
 * count and fields (it sizes the buffers and the bond-list capacity);
 * otherwise PGT_ERR_STATE.  Check with pgt_ik_finish.  From then on the
 * engine's device buffers are held for the enqueued (or captured) launches:
 * a later call that would free, grow or replace them (pgt_ik_set_grid, a
 * pgt_ik_compute_frames needing more room or a larger bond capacity) fails
 * with PGT_ERR_STATE until pgt_ik_release_enqueued; after such a call the
 * graph must be captured again.
 * No counterpart in the reference (whose compute runs one frame at a time on
 * the CPU); it serves the batched, graph-replayed sweep over many frames and
 * systems. */
int pgt_ik_enqueue_frames(pgt_ik_engine *e, int nframes, int64_t n, const double *d_pos,
                          const double *d_diam, const double *d_mass, const double *d_vel,
                          const double *d_disp, unsigned fields, double *d_out_virial,
                          double *d_out_kinetic, double *d_out_elastic);

/* Declare that no captured graph will replay this engine's enqueued launches
 * any more, so calls may again regrow its device buffers.  No counterpart in
 * the reference. */
int pgt_ik_release_enqueued(pgt_ik_engine *e);

/* Wait for the engine's stream and check the work enqueued (or replayed)
 * since the last finish: PGT_ERR_RANGE if a particle had more bonds than the
 * capacity (those results are incomplete: release the enqueued work, run
 * pgt_ik_compute_frames, which grows it, and capture again) or an adaptive
 * integral ran out of workspace.  Sets the counts in
 * pgt_ik_get_stats.  No counterpart in the reference. */
int pgt_ik_finish(pgt_ik_engine *e);

/* Run the engine's work on the given hipStream_t (NULL: the engine's own). */
int pgt_ik_set_stream(pgt_ik_engine *e, void *hip_stream);

/* ---- sweep: many systems' frames, one launch per stage -------------------
 * No counterpart in the reference, whose analyze loop
 * (pyglasstools/__init__.py:119-122) computes one system's frames one call
 * at a time.  A sweep computes the frames of several systems, all 2D or all
 * 3D (each its own engine: box, grid, potential, kernel radius, cell reach;
 * particle counts may differ) with one launch per stage over all of them: cell-list passes (3
 * launches per radix pass of the cell keys, 1 to 3 passes), sorted arrays,
 * bond list, field gather, counters, so a captured graph holds 7 to 13
 * nodes whatever the number of systems.  Each system's fields equal its engine's pgt_ik_compute_frames
 * with the tile gather, bit for bit.  Forms: 2D (positions and grid with
 * z = 0) or 3D, Octic kernel, fixed 5-node rule, FP64 or FP32 bond entries
 * (the same for every system), virial and/or kinetic fields; others give
 * PGT_ERR_UNSUPPORTED.  At create the sweep copies what it needs from the
 * engines (geometry, parameters, grid), so they may change or be destroyed
 * afterwards without affecting it; the device pointers in `systems` are kept
 * and read by every run. */
typedef struct pgt_ik_sweep pgt_ik_sweep;

typedef struct pgt_ik_sweep_system {
    pgt_ik_engine *engine;       /* box, grid and parameters set */
    int32_t nframes;             /* >= 1 */
    int64_t n;                   /* particles per frame, 1 .. 2^26 */
    const double *d_pos;         /* device, nframes * n * 3 (2D: z = 0) */
    const double *d_diam;        /* device, nframes * n */
    const double *d_mass;        /* device, nframes * n, or NULL (unit mass) */
    const double *d_vel;         /* device, nframes * n * 3, or NULL (at rest) */
    double *d_out_virial;        /* device, nframes * G * D * D (fields with virial) */
    double *d_out_kinetic;       /* device, nframes * G * D * D (fields with kinetic) */
} pgt_ik_sweep_system;

/* Create a sweep over nsys (1 .. 16) systems on the current device, launching
 * on hip_stream (NULL: a stream of its own).  fields: PGT_FIELD_VIRIAL and/or
 * PGT_FIELD_KINETIC. */
int pgt_ik_sweep_create(int nsys, const pgt_ik_sweep_system *systems, unsigned fields, void *hip_stream,
                        pgt_ik_sweep **out);
void pgt_ik_sweep_destroy(pgt_ik_sweep *s);

/* Run the sweep and wait for it: sizes the work arrays and grows the bond
 * capacity (running again) as pgt_ik_compute_frames does; sets the counts of
 * pgt_ik_sweep_get_stats, totals over every frame of every system:
 * n_particles and n_grid count particles and grid points summed over frames,
 * gather_ms and total_ms are the gather's and the whole sweep's HIP-event
 * times. */
int pgt_ik_sweep_run(pgt_ik_sweep *s);

/* Enqueue the sweep's launches and return: no synchronisation, allocation,
 * memset or event, so they can be captured into a HIP graph on the sweep's
 * stream.  Needs an earlier pgt_ik_sweep_run (PGT_ERR_STATE otherwise); check
 * with pgt_ik_sweep_finish.  While launches may be replayed the sweep must
 * not be destroyed and pgt_ik_sweep_run must not grow its buffers (it fails
 * with PGT_ERR_STATE while held; pgt_ik_sweep_release declares the graph
 * gone). */
int pgt_ik_sweep_enqueue(pgt_ik_sweep *s);
int pgt_ik_sweep_release(pgt_ik_sweep *s);

/* Wait for the sweep's stream and check the work enqueued (or replayed) since
 * the last finish: PGT_ERR_RANGE if a particle had more bonds than the
 * capacity sized by the last run, or a tile item more than 255 bonds (the
 * results are incomplete).  Sets the counts of pgt_ik_sweep_get_stats. */
int pgt_ik_sweep_finish(pgt_ik_sweep *s);

int pgt_ik_sweep_set_stream(pgt_ik_sweep *s, void *hip_stream);
int pgt_ik_sweep_get_stats(pgt_ik_sweep *s, pgt_ik_stats *out);

/* Launch shape: threads per workgroup of the field gather (128 or 256),
 * frames per launch, grid points per gather lane (1, 2 or 4: neighbouring
 * points that share one read of each particle and bond entry), and cell
 * reach (1: cells of at least the search radius, searched one cell to each
 * side; 2: cells of at least half of it, two to each side); 0 keeps the
 * current choice.  Used by the autotuner.  No counterpart in the reference. */
int pgt_ik_set_tuning(pgt_ik_engine *e, int threads_per_block, int frames_per_launch, int points_per_lane,
                      int cell_reach);

/* Field-gather kernel for the Octic 5-node rule (the driver scripts' rule):
 * 1, the tile gather (one wavefront per 64 grid points lists the (point,
 * particle) pairs of its balls, sorts them by bond count and integrates
 * them round-robin, so every lane has about the same number of bonds);
 * 0, the lane gather (each lane walks the balls of its own one, two or four
 * points).  Same results to rounding; other rules always use the lane
 * gather, as do frames of more than 2^26 particles, and a particle with more
 * than 255 bonds makes the job run again with the lane gather.  No
 * counterpart in the reference. */
int pgt_ik_set_gather_mode(pgt_ik_engine *e, int mode);

int pgt_ik_get_stats(pgt_ik_engine *e, pgt_ik_stats *out);

/* ---- learned bond-integral surrogate --------------------------------------
 * No counterpart in the reference.  A tanh MLP 5 -> 64 -> 64 -> 64 -> 1 stands
 * in for the Octic kernel's 5-node bond integral (FixedPointCGFunc::getBondFunc,
 * CoarseGrainFunction.h:85-88) of a 2D system.  Input row (8 floats, the last
 * three zero): a = (r_i - x)/r_cg, b = (r_j - x)/r_cg, r_cg; output: the
 * unnormalised integral w = sum_k w_k (1 - t_k^2)^2_+ (multiply by the
 * normalisation 15/(8 pi r_cg^2) for getBondFunc's value).  Parameters: FP32,
 * pgt_mlp_nparam() of them, laid out W1[5][64] b1 W2[64][64] b2 W3[64][64] b3
 * W4[64] b4.  All pointers are device pointers; work is enqueued on the given
 * hipStream_t (NULL: the null stream) without host synchronisation. */
int pgt_mlp_nparam(void);
/* floats of workspace pgt_mlp_forward / pgt_mlp_grad need for n samples */
size_t pgt_mlp_work_floats(int64_t n);
/* FP32 forward: y[n] */
int pgt_mlp_forward(const float *d_params, int64_t n, const float *d_x, float *d_y, float *d_work, void *stream);
/* FP32 mean-squared loss against t[n] (loss[0]) and its gradient (nparam floats) */
int pgt_mlp_grad(const float *d_params, int64_t n, const float *d_x, const float *d_t, float *d_grad, float *d_loss,
                 float *d_work, void *stream);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected; step counts from 1) */
int pgt_mlp_adam(float *d_params, float *d_m, float *d_v, const float *d_grad, int step, float lr, void *stream);
/* inference on the matrix cores: BF16 weights and activations, FP32 accumulation */
int pgt_mlp_infer(const float *d_params, int64_t n, const float *d_x, float *d_y, void *stream);

/* Training samples from the engine's frame and grid: every (grid point, bond)
 * pair of IrvingKirkwood::compute's loop (IrvingKirkwood.h:80-107) with a
 * nonzero force, as input rows (d_x, 8 floats each) and its 5-node integral
 * (d_t).  *n_pairs gets the count; with d_x NULL only counts; PGT_ERR_RANGE if
 * max_pairs rows are too few, or if the frame has more than 2^31 - 1 pairs
 * (also refused by pgt_ik_compute_surrogate).  Synchronises the engine's
 * stream. */
int pgt_ik_bond_samples(pgt_ik_engine *e, int64_t max_pairs, float *d_x, float *d_t, int64_t *n_pairs);

/* Virial field of the engine's frame (host output, G x 2 x 2) with the bond
 * integral from the network where the error bound of the pair's class (both
 * ends inside the ball: bound_inside; the segment leaves it: bound_crossing)
 * is at most tol, and from the 5-node rule elsewhere (tol < 0: everywhere).
 * tol and the bounds are absolute errors of one unnormalised bond integral,
 * not a tolerance on the field; the bounds are the caller's (the package
 * calibrates them empirically, as twice the largest error seen on held-out
 * frames), not guarantees.  When neither bound is within tol no pair can
 * take the network's value, and the field comes from the plain gather
 * (the pair list is skipped; n_pairs is then its contribution count).
 * Replaces IrvingKirkwood::compute (IrvingKirkwood.h:74-110) for the virial
 * field in this form; counts in pgt_ik_get_stats (n_pairs, n_network). */
int pgt_ik_compute_surrogate(pgt_ik_engine *e, const float *d_params, double bound_inside, double bound_crossing,
                             double tol, double *virial);

/* Batched evaluators on the GPU, host arrays in and out (n x 3 vectors).
 * pgt_eval_pair_force replaces PairPotential::getPairForceDivR / getPairEnergy /
 * getRcut (PairPotential.h:89-113); out_* may be NULL. */
int pgt_eval_pair_force(int potential_kind, const double *pot_params, int n_pot_params,
                        double scaled_rcut, int64_t n, const double *rij, const double *di,
                        const double *dj, double *out_force_divr, double *out_energy,
                        double *out_rcutsq);

/* mode 0: Delta(x, ri)                 (CoarseGrainFunction.h:78-83, getDeltaFunc)
 * mode 1: Delta(x, ri + s*rij)         (CoarseGrainFunction.h:42-45, getObjFunc)
 * mode 2: bond integral over s in [0,1] (CoarseGrainFunction.h:85-88 fixed,
 *                                        :124-127 adaptive, getBondFunc)
 * s may be NULL unless mode == 1; rij may be NULL for mode 0. */
int pgt_eval_cgfunc(int cg_kind, double cg_rcut, int quad_mode, int quad_order, int adapt_limit,
                    double adapt_epsrel, double adapt_epsabs, int mode, int64_t n,
                    const double *s, const double *x, const double *ri, const double *rij,
                    double *out);

#ifdef __cplusplus
}
#endif

#endif /* PGT_IK_H */
